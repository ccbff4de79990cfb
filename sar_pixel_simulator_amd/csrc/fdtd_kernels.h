// Device-side data layout and kernels of the MI355X FDTD engine.
//
// Layout in HBM (per simulation):
//   fields ex ey ez hx hy hz   float32 [nz][ny][pitch], pitch = nx rounded up
//                               to 16 floats (64-byte rows), x fastest
//   Ca, Cb                     float32 [nz][ny][pitch] (absent for uniform media)
//   CPML psi, 24 arrays        d/u [npml][ny][pitch], l/r [nz][ny][npml],
//                               f/b [nz][npml][pitch]
//   phasor accumulators        float4 (hi.re, hi.im, lo.re, lo.im) per sample,
//                               24 planes in the reference's save order
//   twiddles                   per step: double4 (cos,sin E; cos,sin H) and the
//                               scaled float hi/lo split of the same values
//   beam                       float32 [4][nt_beam][ny][nx] (ex ey hx hy)
//   x-face copies (MAIN)       two float4 per (face, z, y): the DFT's x-face
//                               columns as the split E kernel wrote them
//
// Every kernel reads a per-simulation descriptor (Desc) from device memory,
// so one launch can step many simulations of different sizes (a batch), and
// reads its step counter from device memory, so a captured HIP graph of the
// step sequence replays without parameter updates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fdtd_hip.h"

namespace fdtd {

constexpr int kWave = 64;          // CDNA wavefront width; tiles are 64 cells wide in x
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMaxFaces = 6;

enum FaceKind : int { kFaceZ = 0, kFaceX = 1, kFaceY = 2 };

// psi slots (reference names, cpu/fdtd_sim.py:165-187)
enum Psi : int {
    pExz_d, pEyz_d, pExz_u, pEyz_u, pEyx_l, pEzx_l, pEyx_r, pEzx_r,
    pExy_f, pEzy_f, pExy_b, pEzy_b,
    pHxz_d, pHyz_d, pHxz_u, pHyz_u, pHyx_l, pHzx_l, pHyx_r, pHzx_r,
    pHxy_f, pHzy_f, pHxy_b, pHzy_b,
};
// profile rows of Desc::prof
enum Prof : int { BE_DLF, CE_DLF, BH_DLF, CH_DLF, BE_URB, CE_URB, BH_URB, CH_URB };

struct alignas(16) Desc {
    int nx, ny, nz, npml;
    int tfsf_z, read_out, nt, nt_inc;
    int mode, dft_accum, uni_z, n_src;   // planes z >= uni_z hold (ca_u, cb_u) everywhere
    int nt_beam, nfaces, fused, pad1;    // fused: stepper (0 split; 1, 2, 3 fused)
    long long sy, sz;              // field strides (elements)
    long long du_sz, lr_sz, fb_sz; // psi plane strides
    float Db, dtm0, dx, dte0dx;
    float ca_u, cb_u, pad2, pad3;
    float prof[8][FDTD_MAX_PML];
    float *f[2][6];                // ex ey ez hx hy hz; set t&1 current (fused), else set 0
    const float *ca, *cb;
    float *psi[2][24];             // E psi: set 0 only (set 1 aliases it); H psi: both sets
    const float *beam;             // MAIN: injected beam
    float *rec;                    // INCIDENT: recorded beam
    long long beam_cs;             // beam component stride = nt_beam * ny * nx
    const float *src;              // INCIDENT: hard source [2][n_src][ny][nx]
    long long src_cs;
    float4 *acc;                   // phasor accumulators
    long long acc_off[25];         // start of phasor k (float4 units); [24] = total
    int face_kind[kMaxFaces], face_pos[kMaxFaces];
    int face_na[kMaxFaces], face_nb[kMaxFaces];
    long long face_off[kMaxFaces + 1];
    long long dft_items;           // samples of all faces (+ record items, INCIDENT)
    const double4 *tw64;           // [nt]
    const float4 *tw32;            // [2 * nt]: E then H, scaled by 1/2 and 1/4, hi/lo split
    int *step;                     // [0] step counter, [1] the step k_dft runs
    // MAIN: copies of the x-face columns, [2][nz][ny] records of two float4,
    // (ey, ez, hy(x-1), hy), (hz(x-1), hz, 0, 0) at x = npml + 1 (face 0)
    // and nx - npml - 1 (face 1), written by the split E kernel or the
    // halo stepper
    float4 *tapx;
};

// Launch shape of the stencil kernels.  Each lane owns xvec consecutive x
// cells (1, 2 or 4; 0 = chosen from the grid widths) and marches z over
// zchunk cells (0 = whole column); a tile is (64 xvec) x ty cells.  Split
// stepper: ty in {1, 2, 4}, one row per wavefront of a 64 x ty block.
// Fused stepper: ty in {2, 4, 8} rows walked by one wavefront, four
// wavefronts (tiles) per block.
// Tile shape of the stencil kernels (0 = chosen from the grid; see
// group_build).  The defaults are the split stepper's.
struct Launch {
    int ty = 0;      // rows per block (split) or per wavefront (fused)
    int zchunk = 2;  // z cells marched per tile (0 = whole column)
    int xvec = 0;    // x cells per lane
    int xw = 0;      // split stepper: wavefronts side by side along x per block
    int xcd = 0;     // 0: plain block order, 1: XCD-contiguous, 2: banded,
                     // 3 + b: fine bands of 2^b row blocks per XCD
};

// Block order xcd >= kXcdFine: fine bands of 2^(xcd - kXcdFine) row blocks.
constexpr int kXcdFine = 3;

// Shape word bit of the split and strip kernels: XCD-contiguous block order.
constexpr int kShapeXcd = 4;

// Tile grid of a single simulation's split-kernel launch (ntx > 0), a 3-D
// grid of blocks over z chunks t0 + k of zc planes.  band = 0 (plain
// order): ntx x nyb x (z chunks) blocks, block (i, j, k) covers x tile i of
// tx cells and rows j ty .. j ty + ty - 1.  band > 0 (fine bands): the
// chunk's row blocks form panels of kXcds bands of `band` row blocks, and
// block (i, p, k) of a (kXcds band ntx) x (panels) x (z chunks) grid runs
// on XCD i % kXcds (the dispatcher deals workgroups round-robin), taking
// the (i / kXcds)-th tile, x fastest, of band i % kXcds of panel p.  An
// XCD's consecutive blocks then hold x neighbours and, ntx blocks apart,
// y neighbours, so the rows and columns a tile re-reads of its neighbours
// pass through one L2, while all XCDs stream through the same narrow
// window of rows.  Blocks past the last row block are empty.
// flat = 1 (flat tiles): block b covers cells b tx .. b tx + tx - 1 of the
// plane's row-major (y, x) index; band 0: ntx x 1 x (z chunks) blocks, block
// (b, 0, k) for flat block b; band > 0: the flat blocks are the nyb row
// blocks of one column (ntx = 1), dealt to the XCDs in fine bands as above.
// ntx = 0: tiles come from the group's list.
struct TileGrid {
    int ntx, nyb, tx, ty, zc, nz, t0, band, flat;
};

// Blocks of one z chunk of a tile grid (the launch's x extent times its y extent).
__host__ __device__ inline int tilegrid_gx(const TileGrid &g) { return g.band ? kXcds * g.band * g.ntx : g.ntx; }
__host__ __device__ inline int tilegrid_gy(const TileGrid &g) {
    return g.band ? (g.nyb + kXcds * g.band - 1) / (kXcds * g.band) : g.nyb;
}

// Launch::xw value selecting flat tiles for the split stepper: blocks of 4
// wavefronts, each over 64 consecutive cells of a plane's row-major (y, x)
// index (rows padded to the pitch share wavefronts).  Needs one x cell per
// lane and a row pitch of at least 64 floats.
constexpr int kXwFlat = -1;

// Buffer set holding the fields at step t.
__host__ __device__ inline int current_set(const Desc &d, int t) { return d.fused ? (t & 1) : 0; }

}  // namespace fdtd
