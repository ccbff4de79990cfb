// Stencil, phasor and near-to-far kernels of the MI355X FDTD engine.
//
// Must be compiled with -ffp-contract=off: the field and CPML updates keep
// the reference's one-rounding-per-operation arithmetic
// (fdtdsarsim/cpu/field_updates.py, numba without fastmath) so that results
// are bit-identical to the CPU path on the same inputs.
#include <type_traits>

#include "fdtd_kernels.h"

namespace fdtd {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Map a workgroup id to a tile so that each of the 8 XCDs (which receive
// workgroups round-robin) works on one contiguous range of the tile list:
// neighbouring tiles share halo rows, and those then hit the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int n) {
    const int q = n / kXcds, r = n % kXcds;
    const int xcd = bid % kXcds, k = bid / kXcds;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// Split-kernel launch shape word: bits 0-1 hold log2 of the wavefronts a
// block places side by side along x (the rest of its wavefronts take the
// following rows); bit 2 selects the XCD-contiguous block order.  In plain
// order consecutive blocks, which the dispatcher deals out to the XCDs in
// turn, take consecutive tiles, so the whole device streams through one
// narrow address window.

__device__ __forceinline__ int block_tile(int shape, int n) {
    return (shape & kShapeXcd) ? xcd_tile(blockIdx.x, n) : (int)blockIdx.x;
}

// Neighbour lanes of a wavefront through DPP (wave_shl:1 / wave_shr:1, one
// VALU move, no LDS round trip as a shuffle takes): lane_next(v) is lane
// i + 1's v, lane_prev(v) lane i - 1's; lane 63 and lane 0 get 0.  Called
// where the whole wavefront is active.
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ int load_step(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A simulation's step counter is step[0]; step[1] is the step its k_dft
// runs.  The H-half (or fused) kernel copies step[0] into step[1] from one
// lane of the tile at (0, 0, 0), and k_dft, which reads only step[1], has
// its first block advance step[0]: no launch of a step reads a counter
// another launch of the same step changes, and no separate launch is
// needed to advance it.
__device__ __forceinline__ void note_step(const Desc *__restrict__ d, const int4 &tl, int w, int t) {
    if (tl.y == 0 && tl.z == 0 && (tl.w & 0xffff) == 0 && w == 0 && threadIdx.x == 0)
        __hip_atomic_store(d->step + 1, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Per-cell updates shared by the split and the fused steppers.
// ---------------------------------------------------------------------------

// Field and CPML arrays are addressed as global memory: loads and stores
// then compile to global_load/global_store, which the waitcnt logic tracks
// apart from the scalar loads of the descriptor.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) char gchar;

__device__ __forceinline__ gfloat *gp(float *p) { return (gfloat *)p; }
__device__ __forceinline__ const gfloat *gp(const float *p) { return (const gfloat *)p; }

// Element offsets are 32-bit: every array these kernels index holds fewer
// than 2^30 floats (checked when a simulation is created), so the byte
// offset fits the 32-bit register offset of a global access whose base
// stays in scalar registers (one VGPR per address instead of two).
__device__ __forceinline__ const gfloat *at(const gfloat *b, uint32_t o) {
    return (const gfloat *)((const gchar *)b + (o << 2));
}
__device__ __forceinline__ gfloat *at(gfloat *b, uint32_t o) { return (gfloat *)((gchar *)b + (o << 2)); }

// Scalars of one simulation at one time step, read from the descriptor once
// per kernel so that the cell updates take them from registers.
struct StepK {
    int nx, ny, nz, np;
    int tfsf_z;
    bool inj;                      // MAIN and t < nt_inc: incident-field injection
    bool hard;                     // INCIDENT and t < n_src: hard source
    float Db, dtm0, dx, dte0dx;
    uint32_t sy, sz, du_sz, lr_sz, fb_sz;
    // CPML state: each field's twelve psi arrays lie in one block in the
    // order of enum Psi (d, u: [npml][ny][pitch]; l, r: [nz][ny][npml];
    // f, b: [nz][npml][pitch]), so an array is its block's base plus a
    // closed-form offset (psi_off)
    uint32_t du_len, lr_len, fb_len;
    gfloat *psiE;                      // E block (updated in place)
    const gfloat *b0, *b1, *b2, *b3;   // plane t of the beam (ex ey hx hy) or source (ex ey)
};

// Offset of psi array j (0..11 within one field's block, enum Psi order).
__device__ __forceinline__ uint32_t psi_off(const StepK &k, int j) {
    return j < 4 ? j * k.du_len
         : j < 8 ? 4 * k.du_len + (j - 4) * k.lr_len
                 : 4 * k.du_len + 4 * k.lr_len + (j - 8) * k.fb_len;
}

// The time-independent part of StepK.
__device__ __forceinline__ StepK step_geom(const Desc *__restrict__ d) {
    StepK k;
    k.nx = d->nx; k.ny = d->ny; k.nz = d->nz; k.np = d->npml;
    k.tfsf_z = d->tfsf_z;
    k.inj = k.hard = false;
    k.Db = d->Db; k.dtm0 = d->dtm0; k.dx = d->dx; k.dte0dx = d->dte0dx;
    k.sy = (uint32_t)d->sy; k.sz = (uint32_t)d->sz;
    k.du_sz = (uint32_t)d->du_sz; k.lr_sz = (uint32_t)d->lr_sz; k.fb_sz = (uint32_t)d->fb_sz;
    k.du_len = (uint32_t)(d->npml * d->du_sz);
    k.lr_len = (uint32_t)(d->nz * d->lr_sz);
    k.fb_len = (uint32_t)(d->nz * d->fb_sz);
    k.psiE = gp(d->psi[0][0]);
    k.b0 = k.b1 = k.b2 = k.b3 = nullptr;
    return k;
}

// What the step-dependent part of StepK is formed from.
struct StepTime {
    int nt, nt_inc, n_src, main;
    const gfloat *beam, *src;
    long long beam_cs, src_cs, plane;
};

__device__ __forceinline__ StepTime step_time(const Desc *__restrict__ d) {
    StepTime s;
    s.nt = d->nt; s.nt_inc = d->nt_inc; s.n_src = d->n_src;
    s.main = d->mode == FDTD_MODE_MAIN;
    s.beam = gp(d->beam); s.src = gp(d->src);
    s.beam_cs = d->beam_cs; s.src_cs = d->src_cs;
    s.plane = (long long)d->ny * d->nx;
    return s;
}

// StepK of step t: the injection flags and the beam (or source) plane.
__device__ __forceinline__ StepK step_at(StepK k, const StepTime &s, int t) {
    k.inj = s.main && t < s.nt_inc;
    k.hard = !s.main && t < s.n_src;
    const long long plane = s.plane * t;
    if (k.inj) {
        const gfloat *b = s.beam + plane;
        k.b0 = b;
        k.b1 = b + s.beam_cs;
        k.b2 = b + 2 * s.beam_cs;
        k.b3 = b + 3 * s.beam_cs;
    } else if (k.hard) {
        k.b0 = s.src + plane;
        k.b1 = s.src + s.src_cs + plane;
    }
    return k;
}

__device__ __forceinline__ StepK step_k(const Desc *__restrict__ d, int t) {
    return step_at(step_geom(d), step_time(d), t);
}

// Holds the tile's descriptor scalars in SGPRs at this point, so that their
// loads are issued as one batch ahead of the field loads instead of one
// dependent round trip each after them.
__device__ __forceinline__ void scalars_resident(const StepK &k, const StepTime &s) {
    asm volatile("" ::"s"(k.nx), "s"(k.ny), "s"(k.nz), "s"(k.np), "s"(k.tfsf_z), "s"(k.sy), "s"(k.sz),
                 "s"(k.du_sz), "s"(k.lr_sz), "s"(k.fb_sz), "s"(k.psiE));
    asm volatile("" ::"s"(k.Db), "s"(k.dtm0), "s"(k.dx), "s"(k.dte0dx), "s"(s.nt), "s"(s.nt_inc),
                 "s"(s.n_src), "s"(s.main), "s"(s.beam), "s"(s.src), "s"(s.beam_cs), "s"(s.src_cs),
                 "s"(s.plane));
}

// E values around an H node (x, y, z) at the old time level.
struct EAround {
    float ex_c, ex_n, ex_yp;   // ex(z), ex(z+1), ex(y+1)
    float ey_c, ey_n, ey_xp;   // ey(z), ey(z+1), ey(x+1)
    float ez_c, ez_yp, ez_xp;  // ez(z), ez(y+1), ez(x+1)
};

// What the x-face CPML corrections of one cell read, loaded ahead by the
// split kernels: the two psi values of its l or r face with the cell's
// plane loads, and the face's profile coefficients once per lane (its x is
// fixed as it marches z).  An x-face cell then waits for no dependent load
// after its Yee update.  The loads mirror h_cpml / e_cpml (conditions,
// slots, offsets).  761 x 761 x 203: step -0.6 %; 16 small soil models as
// one batch: -5 %.  The z- and y-face values loaded ahead as well raised
// the register count past 64 (fewer wavefronts per SIMD) and gained less.
struct PsiPre {
    float x[2];
    float xpb, xpc;        // x-face profile coefficients (b, c) of the cell
};

__device__ __forceinline__ void h_xprof_load(PsiPre &q, const Desc *__restrict__ d, const StepK &k, int x) {
    const int nx = k.nx, np = k.np;
    q.xpb = q.xpc = 0.f;
    if (x < np - 1) {
        q.xpb = d->prof[BH_DLF][x];
        q.xpc = d->prof[CH_DLF][x];
    } else if (x >= nx - np && x < nx) {
        q.xpb = d->prof[BH_URB][x - (nx - np)];
        q.xpc = d->prof[CH_URB][x - (nx - np)];
    }
}

__device__ __forceinline__ void e_xprof_load(PsiPre &q, const Desc *__restrict__ d, const StepK &k, int x) {
    const int nx = k.nx, np = k.np;
    q.xpb = q.xpc = 0.f;
    if (x < np) {
        q.xpb = d->prof[BE_DLF][x];
        q.xpc = d->prof[CE_DLF][x];
    } else if (x >= nx - np && x < nx) {
        q.xpb = d->prof[BE_URB][x - (nx - np)];
        q.xpc = d->prof[CE_URB][x - (nx - np)];
    }
}

__device__ __forceinline__ void h_psi_load(PsiPre &q, const StepK &k, const gfloat *hr, int x, int y, int z) {
    q.x[0] = q.x[1] = 0.f;
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    if (z >= nz - 1) return;                                // h_cpml runs below the top plane
    if (yin) {
        if (x < np - 1) {
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)x;
            q.x[0] = *at(hr, psi_off(k, pHzx_l - 12) + p);
            q.x[1] = *at(hr, psi_off(k, pHyx_l - 12) + p);
        }
        if (x >= nx - np && xin) {
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)(x - (nx - np));
            q.x[0] = *at(hr, psi_off(k, pHzx_r - 12) + p);
            q.x[1] = *at(hr, psi_off(k, pHyx_r - 12) + p);
        }
    }
}

__device__ __forceinline__ void e_psi_load(PsiPre &q, const StepK &k, const gfloat *psi, int x, int y, int z) {
    q.x[0] = q.x[1] = 0.f;
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const bool xlo = x >= 1;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    if (z >= nz - 1) return;                                // e_cpml runs below the top plane
    if (yin) {
        if (xlo && x < np) {
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)x;
            q.x[0] = *at(psi, psi_off(k, pEzx_l) + p);
            q.x[1] = *at(psi, psi_off(k, pEyx_l) + p);
        }
        if (x >= nx - np && xin) {
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)(x - (nx - np));
            q.x[0] = *at(psi, psi_off(k, pEzx_r) + p);
            q.x[1] = *at(psi, psi_off(k, pEyx_r) + p);
        }
    }
}

// H half-step of one cell: Yee update (cpu/field_updates.py:198-240), the
// six H-CPML corrections in the reference order d,u,l,r,f,b (:243-366,
// cpu/fdtd_sim.py:104-133), then the E_inc injection at tfsf_z in MAIN mode
// (cpu/main_sim.py:147-150).  Every correction reads E only, so applying
// them in sequence per cell gives the values of the reference's separate
// passes exactly.  psi is read from the H block hr and, when hw is not
// null, the updated value is written to hw (hw == hr updates in place).
// INNER: the caller guarantees the cell lies inside the absorber's inner
// boundary on all six faces (np <= x, y, z and x, y, z < n - np, so also
// away from the grid's outer planes): the update is then the Yee step and
// the injection only, the same operations in the same order.
__device__ __forceinline__ void h_cpml(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       const EAround &e, float &hx, float &hy, float &hz,
                                       const gfloat *hr, gfloat *hw, const PsiPre *pre);

template <bool INNER = false>
__device__ __forceinline__ void h_cell(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       const EAround &e, float &hx, float &hy, float &hz,
                                       const gfloat *hr, gfloat *hw, const PsiPre *pre = nullptr) {
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const float Db = k.Db;
    const bool xin = INNER || x < nx - 1, yin = INNER || y < ny - 1;
    if (INNER || z < nz - 1) {
        if (INNER || (x >= 1 && xin && yin)) hx -= Db * ((e.ez_yp - e.ez_c) - (e.ey_n - e.ey_c));
        if (INNER || (y >= 1 && yin && xin)) hy -= Db * ((e.ex_n - e.ex_c) - (e.ez_xp - e.ez_c));
        if (INNER || (z >= 1 && yin && xin)) hz -= Db * ((e.ey_xp - e.ey_c) - (e.ex_yp - e.ex_c));
        if (!INNER && (z < np - 1 || z >= nz - np || x < np - 1 || x >= nx - np || y < np - 1 || y >= ny - np))
            h_cpml(d, k, x, y, z, e, hx, hy, hz, hr, hw, pre);
    }
    if (k.inj && z == k.tfsf_z) {
        const uint32_t q = (uint32_t)y * nx + x;
        hx += Db * *at(k.b1, q);   // ey_inc
        hy -= Db * *at(k.b0, q);   // ex_inc
    }
}

// The six H-CPML corrections of one cell below the top plane, in the
// reference order d,u,l,r,f,b, after its Yee update.  psi is read from the
// H block hr and, when hw is not null, written to the H block hw (hw == hr
// updates in place).  With pre, the psi values and the x-face profile
// coefficients of the x faces come from the split kernel's loads ahead
// (PsiPre); without, they are loaded here (fused steppers).
__device__ __forceinline__ void h_cpml(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       const EAround &e, float &hx, float &hy, float &hz,
                                       const gfloat *hr, gfloat *hw, const PsiPre *pre) {
    // the cell's x as an opaque value: the absorber's x-dependent profile
    // loads and psi offsets are then formed here, when a cell needs them,
    // instead of being hoisted out of the stencil loops into registers
    asm volatile("" : "+v"(x));
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const uint32_t sy = k.sy;
    const float dtm0 = k.dtm0, dx = k.dx;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    if (yin && xin) {
        if (z < np - 1) {                                       // down
            const float bh = d->prof[BH_DLF][z], ch = d->prof[CH_DLF][z];
            const uint32_t p = (uint32_t)z * k.du_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(hr, psi_off(k, pHyz_d - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ex_n - e.ex_c)) / dx);
            const float pb = *at(hr, psi_off(k, pHxz_d - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ey_n - e.ey_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHyz_d - 12) + p) = a; *at(hw, psi_off(k, pHxz_d - 12) + p) = b; }
            hy -= dtm0 * a;
            hx += dtm0 * b;
        }
        if (z >= nz - np) {                                     // up
            const int zl = z - (nz - np);
            const float bh = d->prof[BH_URB][zl], ch = d->prof[CH_URB][zl];
            const uint32_t p = (uint32_t)zl * k.du_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(hr, psi_off(k, pHyz_u - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ex_n - e.ex_c)) / dx);
            const float pb = *at(hr, psi_off(k, pHxz_u - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ey_n - e.ey_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHyz_u - 12) + p) = a; *at(hw, psi_off(k, pHxz_u - 12) + p) = b; }
            hy -= dtm0 * a;
            hx += dtm0 * b;
        }
    }
    if (yin) {
        if (x < np - 1) {                                       // left
            const float bh = pre ? pre->xpb : d->prof[BH_DLF][x], ch = pre ? pre->xpc : d->prof[CH_DLF][x];
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)x;
            const float pa = pre ? pre->x[0] : *at(hr, psi_off(k, pHzx_l - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ey_xp - e.ey_c)) / dx);
            const float pb = pre ? pre->x[1] : *at(hr, psi_off(k, pHyx_l - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ez_xp - e.ez_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHzx_l - 12) + p) = a; *at(hw, psi_off(k, pHyx_l - 12) + p) = b; }
            hz -= dtm0 * a;
            hy += dtm0 * b;
        }
        if (x >= nx - np && xin) {                              // right
            const int xl = x - (nx - np);
            const float bh = pre ? pre->xpb : d->prof[BH_URB][xl], ch = pre ? pre->xpc : d->prof[CH_URB][xl];
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)xl;
            const float pa = pre ? pre->x[0] : *at(hr, psi_off(k, pHzx_r - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ey_xp - e.ey_c)) / dx);
            const float pb = pre ? pre->x[1] : *at(hr, psi_off(k, pHyx_r - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ez_xp - e.ez_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHzx_r - 12) + p) = a; *at(hw, psi_off(k, pHyx_r - 12) + p) = b; }
            hz -= dtm0 * a;
            hy += dtm0 * b;
        }
    }
    if (xin) {
        if (y < np - 1) {                                       // front
            const float bh = d->prof[BH_DLF][y], ch = d->prof[CH_DLF][y];
            const uint32_t p = (uint32_t)z * k.fb_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(hr, psi_off(k, pHxy_f - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ez_yp - e.ez_c)) / dx);
            const float pb = *at(hr, psi_off(k, pHzy_f - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ex_yp - e.ex_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHxy_f - 12) + p) = a; *at(hw, psi_off(k, pHzy_f - 12) + p) = b; }
            hx -= dtm0 * a;
            hz += dtm0 * b;
        }
        if (y >= ny - np && yin) {                              // back
            const int yl = y - (ny - np);
            const float bh = d->prof[BH_URB][yl], ch = d->prof[CH_URB][yl];
            const uint32_t p = (uint32_t)z * k.fb_sz + (uint32_t)yl * sy + (uint32_t)x;
            const float pa = *at(hr, psi_off(k, pHxy_b - 12) + p);
            const float a = (bh * pa) + ((ch * (e.ez_yp - e.ez_c)) / dx);
            const float pb = *at(hr, psi_off(k, pHzy_b - 12) + p);
            const float b = (bh * pb) + ((ch * (e.ex_yp - e.ex_c)) / dx);
            if (hw) { *at(hw, psi_off(k, pHxy_b - 12) + p) = a; *at(hw, psi_off(k, pHzy_b - 12) + p) = b; }
            hx -= dtm0 * a;
            hz += dtm0 * b;
        }
    }
}

// H values around an E node (x, y, z) at the new time level.
struct HAround {
    float hx_c, hy_c, hz_c;     // at the node
    float hx_zm, hy_zm;         // z-1
    float hz_ym, hx_ym;         // y-1
    float hz_xm, hy_xm;         // x-1
};

// E half-step of one cell: Yee update with the cell's Ca, Cb
// (cpu/field_updates.py:18-63), the six E-CPML corrections d,u,l,r,f,b
// (:66-191, psi updated in place), then the H_inc injection at tfsf_z (MAIN,
// cpu/main_sim.py:158-161) or the hard source at the feed-in plane
// (INCIDENT, cpu/incident_sim.py:142-144).  ca, cb are read only below the
// top plane (z < nz - 1).  INNER as in h_cell.
__device__ __forceinline__ void e_cpml(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       float cb, const HAround &h, float &ex, float &ey, float &ez,
                                       const PsiPre *pre);

template <bool INNER = false>
__device__ __forceinline__ void e_cell(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       float ca, float cb, const HAround &h, float &ex, float &ey,
                                       float &ez, const PsiPre *pre = nullptr) {
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const bool xlo = INNER || x >= 1, ylo = INNER || y >= 1, zlo = INNER || z >= 1;
    const bool xin = INNER || x < nx - 1, yin = INNER || y < ny - 1;
    if (INNER || z < nz - 1) {
        if (zlo && ylo && yin && xin) ex = (ex * ca) + (cb * ((h.hz_c - h.hz_ym) - (h.hy_c - h.hy_zm)));
        if (zlo && yin && xlo && xin) ey = (ey * ca) + (cb * ((h.hx_c - h.hx_zm) - (h.hz_c - h.hz_xm)));
        if (ylo && yin && xlo && xin) ez = (ez * ca) + (cb * ((h.hy_c - h.hy_xm) - (h.hx_c - h.hx_ym)));
        if (!INNER && (z < np || z >= nz - np || x < np || x >= nx - np || y < np || y >= ny - np))
            e_cpml(d, k, x, y, z, cb, h, ex, ey, ez, pre);
    }
    if (z == k.tfsf_z) {
        const uint32_t q = (uint32_t)y * nx + x;
        if (k.inj) {
            ey += k.dte0dx * *at(k.b2, q);   // hx_inc
            ex -= k.dte0dx * *at(k.b3, q);   // hy_inc
        } else if (k.hard && ylo && yin && xin) {
            ex = *at(k.b0, q);
            ey = *at(k.b1, q);
        }
    }
}

// The six E-CPML corrections of one cell below the top plane, in the
// reference order d,u,l,r,f,b (psi updated in place), after its Yee update;
// x-face psi values and profile coefficients from pre as in h_cpml.
__device__ __forceinline__ void e_cpml(const Desc *__restrict__ d, const StepK &k, int x, int y, int z,
                                       float cb, const HAround &h, float &ex, float &ey, float &ez,
                                       const PsiPre *pre) {
    asm volatile("" : "+v"(x));   // as in h_cpml
    const int nx = k.nx, ny = k.ny, nz = k.nz, np = k.np;
    const uint32_t sy = k.sy;
    const float dx = k.dx;
    const bool xlo = x >= 1, ylo = y >= 1, zlo = z >= 1;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    gfloat *psi = k.psiE;
    const float cbdx = cb * dx;
    if (yin && xin) {
        if (zlo && z < np) {                                    // down
            const float be = d->prof[BE_DLF][z], ce = d->prof[CE_DLF][z];
            const uint32_t p = (uint32_t)z * k.du_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(psi, psi_off(k, pEyz_d) + p);
            const float a = (be * pa) + ((ce * (h.hx_c - h.hx_zm)) / dx);
            const float pb = *at(psi, psi_off(k, pExz_d) + p);
            const float b = (be * pb) + ((ce * (h.hy_c - h.hy_zm)) / dx);
            *at(psi, psi_off(k, pEyz_d) + p) = a; *at(psi, psi_off(k, pExz_d) + p) = b;
            ey += cbdx * a;
            ex -= cbdx * b;
        }
        if (z >= nz - np) {                                     // up
            const int zl = z - (nz - np);
            const float be = d->prof[BE_URB][zl], ce = d->prof[CE_URB][zl];
            const uint32_t p = (uint32_t)zl * k.du_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(psi, psi_off(k, pEyz_u) + p);
            const float a = (be * pa) + ((ce * (h.hx_c - h.hx_zm)) / dx);
            const float pb = *at(psi, psi_off(k, pExz_u) + p);
            const float b = (be * pb) + ((ce * (h.hy_c - h.hy_zm)) / dx);
            *at(psi, psi_off(k, pEyz_u) + p) = a; *at(psi, psi_off(k, pExz_u) + p) = b;
            ey += cbdx * a;
            ex -= cbdx * b;
        }
    }
    if (yin) {
        if (xlo && x < np) {                                    // left
            const float be = pre ? pre->xpb : d->prof[BE_DLF][x], ce = pre ? pre->xpc : d->prof[CE_DLF][x];
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)x;
            const float pa = pre ? pre->x[0] : *at(psi, psi_off(k, pEzx_l) + p);
            const float a = (be * pa) + ((ce * (h.hy_c - h.hy_xm)) / dx);
            const float pb = pre ? pre->x[1] : *at(psi, psi_off(k, pEyx_l) + p);
            const float b = (be * pb) + ((ce * (h.hz_c - h.hz_xm)) / dx);
            *at(psi, psi_off(k, pEzx_l) + p) = a; *at(psi, psi_off(k, pEyx_l) + p) = b;
            ez += cbdx * a;
            ey -= cbdx * b;
        }
        if (x >= nx - np && xin) {                              // right
            const int xl = x - (nx - np);
            const float be = pre ? pre->xpb : d->prof[BE_URB][xl], ce = pre ? pre->xpc : d->prof[CE_URB][xl];
            const uint32_t p = (uint32_t)z * k.lr_sz + (uint32_t)y * np + (uint32_t)xl;
            const float pa = pre ? pre->x[0] : *at(psi, psi_off(k, pEzx_r) + p);
            const float a = (be * pa) + ((ce * (h.hy_c - h.hy_xm)) / dx);
            const float pb = pre ? pre->x[1] : *at(psi, psi_off(k, pEyx_r) + p);
            const float b = (be * pb) + ((ce * (h.hz_c - h.hz_xm)) / dx);
            *at(psi, psi_off(k, pEzx_r) + p) = a; *at(psi, psi_off(k, pEyx_r) + p) = b;
            ez += cbdx * a;
            ey -= cbdx * b;
        }
    }
    if (xin) {
        if (ylo && y < np) {                                    // front
            const float be = d->prof[BE_DLF][y], ce = d->prof[CE_DLF][y];
            const uint32_t p = (uint32_t)z * k.fb_sz + (uint32_t)y * sy + (uint32_t)x;
            const float pa = *at(psi, psi_off(k, pExy_f) + p);
            const float a = (be * pa) + ((ce * (h.hz_c - h.hz_ym)) / dx);
            const float pb = *at(psi, psi_off(k, pEzy_f) + p);
            const float b = (be * pb) + ((ce * (h.hx_c - h.hx_ym)) / dx);
            *at(psi, psi_off(k, pExy_f) + p) = a; *at(psi, psi_off(k, pEzy_f) + p) = b;
            ex += cbdx * a;
            ez -= cbdx * b;
        }
        if (y >= ny - np && yin) {                              // back
            const int yl = y - (ny - np);
            const float be = d->prof[BE_URB][yl], ce = d->prof[CE_URB][yl];
            const uint32_t p = (uint32_t)z * k.fb_sz + (uint32_t)yl * sy + (uint32_t)x;
            const float pa = *at(psi, psi_off(k, pExy_b) + p);
            const float a = (be * pa) + ((ce * (h.hz_c - h.hz_ym)) / dx);
            const float pb = *at(psi, psi_off(k, pEzy_b) + p);
            const float b = (be * pb) + ((ce * (h.hx_c - h.hx_ym)) / dx);
            *at(psi, psi_off(k, pExy_b) + p) = a; *at(psi, psi_off(k, pEzy_b) + p) = b;
            ex += cbdx * a;
            ez -= cbdx * b;
        }
    }
}

// ---------------------------------------------------------------------------
// Lane vectors: each lane owns V consecutive x cells of one row and loads
// and stores them as one V-wide access.  Rows are padded to a multiple of
// 16 floats, so a lane's V cells are either all inside the row or all in
// its padding.
// ---------------------------------------------------------------------------
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vf2 gvf2;
typedef __attribute__((address_space(1))) vf4 gvf4;

template <int V>
__device__ __forceinline__ void vld(float (&r)[V], const gfloat *p) {
    if constexpr (V == 4) {
        const vf4 q = *reinterpret_cast<const gvf4 *>(p);
        r[0] = q.x; r[1] = q.y; r[2] = q.z; r[3] = q.w;
    } else if constexpr (V == 2) {
        const vf2 q = *reinterpret_cast<const gvf2 *>(p);
        r[0] = q.x; r[1] = q.y;
    } else {
        r[0] = *p;
    }
}

// Non-temporal ("streaming") loads and stores.  At 761 x 761 x 203 the
// split kernels run fastest with these streamed: the material coefficients
// Ca, Cb (read once per step, never written while stepping), the E
// half-step's loads of its own ex, ey, ez, both half-steps' field stores,
// and the DFT's running sums.  Streaming the other loads (the H half-step's
// own fields, the E half-step's hy, ey in the H half-step, the neighbour
// rows) was measured slower.  Together: E half 0.941 -> 0.900 ms, step
// 1.820 -> 1.765 ms.
template <int V>
__device__ __forceinline__ void vld_stream(float (&r)[V], const gfloat *p) {
    if constexpr (V == 4) {
        const vf4 q = __builtin_nontemporal_load(reinterpret_cast<const gvf4 *>(p));
        r[0] = q.x; r[1] = q.y; r[2] = q.z; r[3] = q.w;
    } else if constexpr (V == 2) {
        const vf2 q = __builtin_nontemporal_load(reinterpret_cast<const gvf2 *>(p));
        r[0] = q.x; r[1] = q.y;
    } else {
        r[0] = __builtin_nontemporal_load(p);
    }
}

template <int V>
__device__ __forceinline__ void vst(gfloat *p, const float (&r)[V]) {
    if constexpr (V == 4) {
        *reinterpret_cast<gvf4 *>(p) = vf4{r[0], r[1], r[2], r[3]};
    } else if constexpr (V == 2) {
        *reinterpret_cast<gvf2 *>(p) = vf2{r[0], r[1]};
    } else {
        *p = r[0];
    }
}

template <int V>
__device__ __forceinline__ void vst_stream(gfloat *p, const float (&r)[V]) {
    if constexpr (V == 4) {
        __builtin_nontemporal_store(vf4{r[0], r[1], r[2], r[3]}, reinterpret_cast<gvf4 *>(p));
    } else if constexpr (V == 2) {
        __builtin_nontemporal_store(vf2{r[0], r[1]}, reinterpret_cast<gvf2 *>(p));
    } else {
        __builtin_nontemporal_store(r[0], p);
    }
}

template <int V>
__device__ __forceinline__ void vset(float (&r)[V], float v) {
#pragma unroll
    for (int k = 0; k < V; ++k) r[k] = v;
}

template <int V>
__device__ __forceinline__ void vcopy(float (&r)[V], const float (&s)[V]) {
#pragma unroll
    for (int k = 0; k < V; ++k) r[k] = s[k];
}

// ---------------------------------------------------------------------------
// Split stepper: H half-step kernel, then E half-step kernel, fields updated
// in place (set 0).  Used for INCIDENT runs (E before H) and available for
// MAIN runs.  A block of 64 x (xw ty) lanes covers a (64 V xw) x ty tile:
// xw wavefronts side by side along each of ty rows (xw ty <= 4).  Wide
// blocks (xw = 4: 1 KB per array per block at V = 1) in plain block order
// keep the device's concurrent accesses within a narrow window of each
// array, which the DRAM serves faster than accesses spread over many rows
// and planes.  Each lane marches z upward through the tile's z range,
// reusing the plane above (H half: ex, ey) or below (E half: hx, hy) from
// registers.  (Issuing plane z + 1's loads before computing plane z was
// measured slower: the extra registers cost more occupancy than it hides.)
// The x +- 1 neighbours come from the lane's own cells or, at its ends, from
// the adjacent lane by a DPP lane move (from memory at the tile edge).
// ---------------------------------------------------------------------------

// What the H half-step loads per plane: H(z), E(z+1) (ex, ey), ez(z),
// ez(y+1), ex(y+1); lane 63: ey, ez at the next tile's first cell.
template <int V>
struct HPlane {
    float hx[V], hy[V], hz[V], exn[V], eyn[V], ez[V], ezyp[V], exyp[V];
    float ey_x, ez_x;
};

template <int V>
__device__ __forceinline__ void h_plane_load(HPlane<V> &P, const gfloat *Ex, const gfloat *Ey,
                                             const gfloat *Ez, const gfloat *Hx, const gfloat *Hy,
                                             const gfloat *Hz, uint32_t o, uint32_t sy, uint32_t sz,
                                             bool act, bool zin, bool yin, bool next_mem) {
    vset(P.exn, 0.f); vset(P.eyn, 0.f); vset(P.ez, 0.f); vset(P.ezyp, 0.f); vset(P.exyp, 0.f);
    vset(P.hx, 0.f); vset(P.hy, 0.f); vset(P.hz, 0.f);
    P.ey_x = P.ez_x = 0.f;
    if (act) {
        vld(P.hx, at(Hx, o));
        vld(P.hy, at(Hy, o));
        vld(P.hz, at(Hz, o));
        if (zin) {
            vld(P.exn, at(Ex, o + sz));
            vld(P.eyn, at(Ey, o + sz));
            vld(P.ez, at(Ez, o));
            if (yin) {
                vld(P.ezyp, at(Ez, o + sy));
                vld(P.exyp, at(Ex, o + sy));
            }
            if (next_mem) {
                P.ey_x = *at(Ey, o + V);
                P.ez_x = *at(Ez, o + V);
            }
        }
    }
}

// x-face taps.  The DFT's x faces lie at x = np + 1 and x = nx - np - 1;
// their samples need ey, ez at the face column and hy, hz at it and at the
// column below, for every (z, y).  Read from the field arrays these are
// one cache line per value; the E half-step kernel, which holds all six
// (the new H at x and x - 1 are its curl operands), therefore copies them
// into Desc::tapx, where the DFT reads them with unit stride in y.  Only
// the wavefronts whose cells include a face column take the branch.  The
// copies hold every (z, y) with z < nz - 1, all the DFT reads.
template <int V>
__device__ __forceinline__ bool wave_taps(int x0, int nx, int np) {
    const int x1 = x0 + kWave * V;
    const int p0 = np + 1, p1 = nx - np - 1;
    return (p0 >= x0 && p0 < x1) || (p1 >= x0 && p1 < x1);
}

// hy_xm, hz_xm: the new hy, hz at the x - 1 neighbour of the lane's first
// cell.  One 32-byte record per (face, z, y): Desc::tapx.
template <int V>
__device__ __forceinline__ void e_tap(gchar *tap, int xb, int nx, int np, int nz, int ny, int z, int y,
                                      const float (&ey)[V], const float (&ez)[V], const float (&hy)[V],
                                      const float (&hz)[V], float hy_xm, float hz_xm) {
#pragma unroll
    for (int c = 0; c < V; ++c) {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            if (xb + c == (f ? nx - np - 1 : np + 1)) {
                const uint32_t cell = (uint32_t)((f * nz + z) * ny + y);
                gvf4 *p = (gvf4 *)(tap + (cell << 5));
                const float hy0 = c > 0 ? hy[c > 0 ? c - 1 : 0] : hy_xm;
                const float hz0 = c > 0 ? hz[c > 0 ? c - 1 : 0] : hz_xm;
                p[0] = vf4{ey[c], ez[c], hy0, hy[c]};
                p[1] = vf4{hz0, hz[c], 0.f, 0.f};
            }
        }
    }
}

// Cell of a split-kernel lane (one x cell per lane in flat tiles).  Row
// tiles: the wavefront's row is y = tl.z + w / xw (uniform), its x range
// starts at the tile's x0.  Flat tiles (FLAT): wavefront w of the block
// covers the 64 consecutive cells fw .. fw + 63 of the plane's row-major
// (y, x) index, fw = tl.y + 64 w; rows are padded to the pitch sy >= 64, so
// a wavefront spans at most two rows, and the lanes past the row's last
// cell (x >= nx) or past the last row (y >= ny) hold no cell.  Narrow rows
// then share wavefronts instead of each being padded to whole ones.
// Returns false when the whole wavefront lies past the grid.
template <int V, bool FLAT>
__device__ __forceinline__ bool lane_cell(const int4 &tl, int w, int lxw, int ny, uint32_t sy, int &y, int &xb) {
    static_assert(V == 1 || !FLAT, "flat tiles hold one x cell per lane");
    const int lane = threadIdx.x;
    if (FLAT) {
        const int fw = tl.y + (w << 6);
        const int yw = __builtin_amdgcn_readfirstlane(fw / (int)sy);
        if (yw >= ny) return false;                    // uniform per wavefront
        y = yw;
        xb = fw - yw * (int)sy + lane;
        if (xb >= (int)sy) {
            xb -= (int)sy;
            ++y;
        }
        return true;
    }
    y = tl.z + (w >> lxw);
    if (y >= ny) return false;                         // uniform per wavefront (one row)
    xb = tl.y + V * (((w & ((1 << lxw) - 1)) << 6) + lane);
    return true;
}

// Tile of a split-kernel block: for a single simulation (DIRECT), from
// the 3-D grid (x tile, row block, z chunk: x fastest, as the plain order;
// no dependent load); else from the group's tile list.  Returns false for a
// padding tile.
template <bool DIRECT>
__device__ __forceinline__ bool split_tile(const int4 *__restrict__ tiles, int ntiles, int shape,
                                           const TileGrid &tg, int4 &tl) {
    if (DIRECT) {
        const int z0 = (int)(blockIdx.z + tg.t0) * tg.zc, z1 = min(tg.nz, z0 + tg.zc);
        int xb = (int)blockIdx.x, yb = (int)blockIdx.y;
        if (tg.band) {
            // fine bands: slot i of panel p; XCD i % 8 takes the
            // (i / 8)-th tile of its band, x fastest
            const int k = xb / kXcds, q = k / tg.ntx;
            yb = (yb * kXcds + xb % kXcds) * tg.band + q;
            xb = k - q * tg.ntx;
            if (yb >= tg.nyb) return false;
        }
        if (tg.flat) {
            // flat tiles: block b covers the plane's cells from b tx on; in
            // fine bands the flat blocks are the row blocks of one column
            tl = make_int4(0, (tg.band ? yb : xb) * tg.tx, 0, z0 | (z1 << 16));
            return true;
        }
        tl = make_int4(0, xb * tg.tx, yb * tg.ty, z0 | (z1 << 16));
        return true;
    }
    tl = tiles[block_tile(shape, ntiles)];
    return tl.w != 0;                                  // padding tile of the banded order
}

// Both split kernels read the descriptor's scalars and, for a single
// simulation, the step counter in one batch of scalar loads, then issue the
// tile's first plane loads (and its x-face CPML values), and test the step
// after that; no field store precedes the test (the H kernel's one-lane copy
// of the counter for k_dft does, note_step).  The z loop then computes a
// plane, stores it and loads the next one.
template <int V, bool DIRECT, bool FLAT>
__global__ void __launch_bounds__(256)
k_h_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles, int shape, TileGrid tg,
         const int *__restrict__ stepp) {
    int4 tl;
    if (!split_tile<DIRECT>(tiles, ntiles, shape, tg, tl)) return;
    const Desc *__restrict__ d = descs + (tl.x & 0xffff);
    // a single simulation's counter is a plain load through a read-only
    // kernel argument, a scalar load in the descriptor's batch: the
    // previous kernel of the stream wrote it, and nothing writes it now
    const int t = DIRECT ? *stepp : load_step(d->step);
    const StepK kg = step_geom(d);
    const StepTime st = step_time(d);
    scalars_resident(kg, st);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    note_step(d, tl, w, t);
    const int lxw = DIRECT ? shape & 3 : (tl.x >> 16) & 3;   // listed tiles carry their block width
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const uint32_t sy = (uint32_t)d->sy, sz = (uint32_t)d->sz;
    int y, xb;                                         // the lane's row and first cell
    if (!lane_cell<V, FLAT>(tl, w, lxw, ny, sy, y, xb)) return;
    const int lane = threadIdx.x;
    const bool act = xb < nx && (!FLAT || y < ny);
    // the x + 1 neighbour of the lane's last cell is needed when that cell
    // is below nx - 1; for lane 63 it lies in the next tile
    const bool next_mem = lane == kWave - 1 && xb + V < nx && act;
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const gfloat *Ex = gp(d->f[0][0]), *Ey = gp(d->f[0][1]), *Ez = gp(d->f[0][2]);
    gfloat *Hx = gp(d->f[0][3]), *Hy = gp(d->f[0][4]), *Hz = gp(d->f[0][5]);
    gfloat *psiH = gp(d->psi[0][pHxz_d]);
    const bool yin = y < ny - 1;

    uint32_t o = (uint32_t)z0 * sz + (uint32_t)y * sy + (uint32_t)xb;
    float exc[V], eyc[V];
    vset(exc, 0.f);
    vset(eyc, 0.f);
    if (act) {
        vld(exc, at(Ex, o));
        vld(eyc, at(Ey, o));
    }
    HPlane<V> P;
    h_plane_load(P, Ex, Ey, Ez, Hx, Hy, Hz, o, sy, sz, act, z0 < nz - 1, yin, next_mem);
    PsiPre pq[V] = {};
#pragma unroll
    for (int c = 0; c < V; ++c) {
        h_xprof_load(pq[c], d, kg, xb + c);
        if (xb + c < nx && act) h_psi_load(pq[c], kg, psiH, xb + c, y, z0);
    }
    if (t >= st.nt) return;                            // uniform per block
    const StepK k = step_at(kg, st, t);
    for (int z = z0;;) {
        float ey_next = lane_next(eyc[0]);
        float ez_next = lane_next(P.ez[0]);
        if (lane == kWave - 1) {
            ey_next = P.ey_x;
            ez_next = P.ez_x;
        }
#pragma unroll
        for (int c = 0; c < V; ++c) {
            if (xb + c < k.nx && act) {
                EAround e;
                e.ex_c = exc[c]; e.ex_n = P.exn[c]; e.ex_yp = P.exyp[c];
                e.ey_c = eyc[c]; e.ey_n = P.eyn[c];
                e.ey_xp = c + 1 < V ? eyc[c + 1 < V ? c + 1 : 0] : ey_next;
                e.ez_c = P.ez[c]; e.ez_yp = P.ezyp[c];
                e.ez_xp = c + 1 < V ? P.ez[c + 1 < V ? c + 1 : 0] : ez_next;
                h_cell(d, k, xb + c, y, z, e, P.hx[c], P.hy[c], P.hz[c], psiH, psiH, &pq[c]);
            }
        }
        if (act) {
            vst_stream(at(Hx, o), P.hx);
            vst_stream(at(Hy, o), P.hy);
            vst_stream(at(Hz, o), P.hz);
        }
        if (++z >= z1) break;
        o += sz;
        vcopy(exc, P.exn);
        vcopy(eyc, P.eyn);
        h_plane_load(P, Ex, Ey, Ez, Hx, Hy, Hz, o, sy, sz, act, z < nz - 1, yin, next_mem);
#pragma unroll
        for (int c = 0; c < V; ++c)
            if (xb + c < nx && act) h_psi_load(pq[c], k, psiH, xb + c, y, z);
    }
}

// What the E half-step loads per plane: H(z) (hx, hy, hz), hz(y-1),
// hx(y-1), E(z), Ca/Cb; lane 0: hz, hy at the previous tile's last cell.
template <int V>
struct EPlane {
    float hx[V], hy[V], hz[V], hzym[V], hxym[V], ex[V], ey[V], ez[V], ca[V], cb[V];
    float hz_p, hy_p;
};

template <int V>
__device__ __forceinline__ void e_plane_load(EPlane<V> &Q, const gfloat *Ex, const gfloat *Ey,
                                             const gfloat *Ez, const gfloat *Hx, const gfloat *Hy,
                                             const gfloat *Hz, const gfloat *Ca, const gfloat *Cb,
                                             float ca_u, float cb_u, uint32_t o, uint32_t sy, bool act,
                                             bool zin, bool mat, bool ylo, bool prev_mem) {
    vset(Q.hx, 0.f); vset(Q.hy, 0.f); vset(Q.hz, 0.f); vset(Q.hzym, 0.f); vset(Q.hxym, 0.f);
    vset(Q.ex, 0.f); vset(Q.ey, 0.f); vset(Q.ez, 0.f);
    vset(Q.ca, ca_u);
    vset(Q.cb, cb_u);
    Q.hz_p = Q.hy_p = 0.f;
    if (act) {
        vld(Q.hx, at(Hx, o));
        vld(Q.hy, at(Hy, o));
        vld_stream(Q.ex, at(Ex, o));
        vld_stream(Q.ey, at(Ey, o));
        vld_stream(Q.ez, at(Ez, o));
        if (zin) {
            vld(Q.hz, at(Hz, o));
            if (ylo) {
                vld(Q.hzym, at(Hz, o - sy));
                vld(Q.hxym, at(Hx, o - sy));
            }
            if (mat) {
                vld_stream(Q.ca, at(Ca, o));
                vld_stream(Q.cb, at(Cb, o));
            }
            if (prev_mem) {
                Q.hz_p = *at(Hz, o - 1);
                Q.hy_p = *at(Hy, o - 1);
            }
        }
    }
}

template <int V, bool DIRECT, bool FLAT>
__global__ void __launch_bounds__(256)
k_e_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles, int shape, TileGrid tg,
         const int *__restrict__ stepp) {
    int4 tl;
    if (!split_tile<DIRECT>(tiles, ntiles, shape, tg, tl)) return;
    const Desc *__restrict__ d = descs + (tl.x & 0xffff);
    // a single simulation's counter is a plain load through a read-only
    // kernel argument, a scalar load in the descriptor's batch: the
    // previous kernel of the stream wrote it, and nothing writes it now
    const int t = DIRECT ? *stepp : load_step(d->step);
    const StepK kg = step_geom(d);
    const StepTime st = step_time(d);
    scalars_resident(kg, st);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lxw = DIRECT ? shape & 3 : (tl.x >> 16) & 3;   // listed tiles carry their block width
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const uint32_t sy = (uint32_t)d->sy, sz = (uint32_t)d->sz;
    int y, xb;                                         // the lane's row and first cell
    if (!lane_cell<V, FLAT>(tl, w, lxw, ny, sy, y, xb)) return;
    const int lane = threadIdx.x;
    const bool act = xb < nx && (!FLAT || y < ny);
    const bool prev_mem = lane == 0 && xb >= 1 && act;    // x - 1 lies in the previous tile
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    gfloat *Ex = gp(d->f[0][0]), *Ey = gp(d->f[0][1]), *Ez = gp(d->f[0][2]);
    const gfloat *Hx = gp(d->f[0][3]), *Hy = gp(d->f[0][4]), *Hz = gp(d->f[0][5]);
    const gfloat *Ca = gp(d->ca), *Cb = gp(d->cb);
    const bool ylo = y >= 1;
    const int uni_z = d->uni_z;
    // the wavefronts whose cells include a DFT x-face column copy it
    bool taps;
    if (FLAT) {
        const int np = d->npml;
        taps = __ballot(act && (xb == np + 1 || xb == nx - np - 1)) != 0;
    } else {
        const int x0w = tl.y + V * ((w & ((1 << lxw) - 1)) << 6);   // the wavefront's first cell
        taps = wave_taps<V>(x0w, nx, d->npml);
    }
    gchar *const tap = d->tapx && taps ? (gchar *)d->tapx : nullptr;
    const float ca_u = d->ca_u, cb_u = d->cb_u;

    uint32_t o = (uint32_t)z0 * sz + (uint32_t)y * sy + (uint32_t)xb;
    float hxzm[V], hyzm[V];
    vset(hxzm, 0.f);
    vset(hyzm, 0.f);
    if (act && z0 >= 1) {
        vld(hxzm, at(Hx, o - sz));
        vld(hyzm, at(Hy, o - sz));
    }
    EPlane<V> P;
    e_plane_load(P, Ex, Ey, Ez, Hx, Hy, Hz, Ca, Cb, ca_u, cb_u, o, sy, act, z0 < nz - 1,
                 z0 < nz - 1 && z0 < uni_z, ylo, prev_mem);
    PsiPre pq[V] = {};
#pragma unroll
    for (int c = 0; c < V; ++c) {
        e_xprof_load(pq[c], d, kg, xb + c);
        if (xb + c < nx && act) e_psi_load(pq[c], kg, kg.psiE, xb + c, y, z0);
    }
    if (t >= st.nt) return;                            // uniform per block
    const StepK k = step_at(kg, st, t);
    for (int z = z0;;) {
        float hz_prev = lane_prev(P.hz[V - 1]);
        float hy_prev = lane_prev(P.hy[V - 1]);
        if (lane == 0) {
            hz_prev = P.hz_p;
            hy_prev = P.hy_p;
        }
#pragma unroll
        for (int c = 0; c < V; ++c) {
            if (xb + c < k.nx && act) {
                HAround h;
                h.hx_c = P.hx[c]; h.hy_c = P.hy[c]; h.hz_c = P.hz[c];
                h.hx_zm = hxzm[c]; h.hy_zm = hyzm[c];
                h.hz_ym = P.hzym[c]; h.hx_ym = P.hxym[c];
                h.hz_xm = c > 0 ? P.hz[c > 0 ? c - 1 : 0] : hz_prev;
                h.hy_xm = c > 0 ? P.hy[c > 0 ? c - 1 : 0] : hy_prev;
                e_cell(d, k, xb + c, y, z, P.ca[c], P.cb[c], h, P.ex[c], P.ey[c], P.ez[c], &pq[c]);
            }
        }
        if (act) {
            vst_stream(at(Ex, o), P.ex);
            vst_stream(at(Ey, o), P.ey);
            vst_stream(at(Ez, o), P.ez);
        }
        if (tap && act) e_tap<V>(tap, xb, nx, k.np, nz, ny, z, y, P.ey, P.ez, P.hy, P.hz, hy_prev, hz_prev);
        if (++z >= z1) break;
        o += sz;
        vcopy(hxzm, P.hx);
        vcopy(hyzm, P.hy);
        e_plane_load(P, Ex, Ey, Ez, Hx, Hy, Hz, Ca, Cb, ca_u, cb_u, o, sy, act, z < nz - 1,
                     z < nz - 1 && z < uni_z, ylo, prev_mem);
#pragma unroll
        for (int c = 0; c < V; ++c)
            if (xb + c < nx && act) e_psi_load(pq[c], k, k.psiE, xb + c, y, z);
    }
}

template __global__ void k_h_half<1, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<1, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<2, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<2, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<4, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<4, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<1, false, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_h_half<1, true, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<1, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<1, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<2, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<2, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<4, false, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<4, true, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<1, false, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_e_half<1, true, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);

// ---------------------------------------------------------------------------
// Fused stepper (MAIN): one kernel per time step computes H(t+1/2) and then
// E(t+1), so each field value is read once and written once per step
// (56 B/cell instead of the split stepper's 80).  Fields and H-CPML state
// are double-buffered: step t reads set t&1 and writes set 1-(t&1).
//
// Each wavefront owns a strip of R rows of 64 V cells over a z range and
// marches z; within a plane it walks its rows upward, computing H(t+1/2) of
// a row and then E(t+1) of the row.  What one plane hands to the next stays
// on chip, in the wavefront's own slice of LDS (no other wavefront touches
// it, so no barrier is needed):
//   - E(t) at z+1 of every row (the row's E(t) at z on the next plane);
//   - H(t+1/2) at z of the strip's rows (H at z-1 for E on the next plane).
// Within a plane, H(t+1/2) of the row below and E(t) ez of the row above
// pass from row to row in registers.  Halo values are recomputed from the
// old level, bit-identical to their owner's, and never stored:
//   - row y0-1 (H at y-1 for the strip's first row), on every plane;
//   - plane z0-1 (H at z-1 for the first plane of a strip with z0 > 0);
//   - cell x0-1 (H at x-1 for lane 0), by lane 0.
// ---------------------------------------------------------------------------

// LDS floats per wavefront: slots of 64 V floats (one value per lane cell):
// ex, ey for rows 0..R (row 0 is y0-1), then hx, hy for rows 1..R.
template <int V, int R>
struct StripLds {
    static constexpr int kSlot = kWave * V;
    static constexpr int kFloats = (4 * R + 2) * kSlot;
    __device__ static int ex(int j) { return (2 * j) * kSlot; }
    __device__ static int ey(int j) { return (2 * j + 1) * kSlot; }
    __device__ static int hx(int j) { return (2 * (R + 1) + 2 * (j - 1)) * kSlot; }
    __device__ static int hy(int j) { return (2 * (R + 1) + 2 * (j - 1) + 1) * kSlot; }
};

template <int V>
__device__ __forceinline__ void lds_ld(float (&r)[V], const float *p) {
#pragma unroll
    for (int c = 0; c < V; ++c) r[c] = p[c];
}

template <int V>
__device__ __forceinline__ void lds_st(float *p, const float (&r)[V]) {
#pragma unroll
    for (int c = 0; c < V; ++c) p[c] = r[c];
}

template <int V, int R>
__global__ void __launch_bounds__(256)
k_step_strip(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles, int shape) {
    using L = StripLds<V, R>;
    __shared__ float lds_all[4][L::kFloats];
    const int nblk = (ntiles + 3) >> 2;
    // one tile per wavefront: threadIdx.y is uniform across the wavefront,
    // and saying so keeps the descriptor and its pointers in scalar registers
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int wt = __builtin_amdgcn_readfirstlane((block_tile(shape, nblk) << 2) + w);
    if (wt >= ntiles) return;                          // uniform per wavefront
    const int4 tl = tiles[wt];
    const Desc *__restrict__ d = descs + (tl.x & 0xffff);   // low 16 bits: member index
    const int t = load_step(d->step);
    note_step(d, tl, 0, t);                            // tiles are per wavefront
    if (t >= d->nt) return;
    const StepK k = step_k(d, t);
    const int lane = threadIdx.x;
    float *lds = lds_all[w] + V * lane;                // this lane's cells of each slot
    const int x0 = tl.y, y0 = tl.z;
    const int xb = x0 + V * lane;
    const bool act = xb < k.nx;
    const bool next_mem = lane == kWave - 1 && xb + V < k.nx;
    const bool halo_x = lane == 0 && x0 > 0;           // recompute H(t+1/2) at x0 - 1
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const int zs = z0 > 0 ? z0 - 1 : 0;                // plane z0 - 1: recomputed, not stored
    const int ry = min(R, k.ny - y0);                  // rows of the strip
    const int p = t & 1;
    const uint32_t sy = k.sy, sz = k.sz;
    const gfloat *fo[6];
    gfloat *fn[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        fo[c] = gp(d->f[p][c]);
        fn[c] = gp(d->f[1 - p][c]);
    }
    const gfloat *psi_o = gp(d->psi[p][pHxz_d]);      // H-CPML state of this step
    gfloat *psi_n = gp(d->psi[1 - p][pHxz_d]);        // and of the next
    const gfloat *Ca = gp(d->ca), *Cb = gp(d->cb);
    const int uni_z = d->uni_z;
    const float ca_u = d->ca_u, cb_u = d->cb_u;

#pragma unroll 1
    for (int z = zs; z < z1; ++z) {
        const bool zin = z < k.nz - 1;
        const bool hplane = z < z0;
        const uint32_t oz = (uint32_t)z * sz + (uint32_t)xb;
        const int j0 = (!hplane && y0 >= 1) ? 0 : 1;   // row 0 (y0 - 1) only where E is updated
        float hzb[V], hxb[V];                          // H(t+1/2) of the row below
        float ezc[V];                                  // E(t) ez of this row, from the row below
        vset(hzb, 0.f);
        vset(hxb, 0.f);
        vset(ezc, 0.f);
#pragma unroll 1
        for (int j = j0; j <= ry; ++j) {
            const int y = y0 - 1 + j;
            const uint32_t o = oz + (uint32_t)y * sy;
            const bool own = j >= 1 && !hplane;
            const bool yin = y < k.ny - 1;
            // ---- every load of the row first, so the row costs one round trip
            // E(t) of the row at z: from LDS where this row was stepped on the
            // plane below, else from memory; ex of the row above likewise
            const bool have = j >= 1 ? z > zs : z > z0;
            const bool have_up = j < ry && z > zs;
            float ex[V], ey[V], exn[V], eyn[V], exyp[V], ezyp[V], hx[V], hy[V], hz[V];
            float hxzm[V], hyzm[V], ca[V], cb[V];
            vset(ex, 0.f); vset(ey, 0.f); vset(exn, 0.f); vset(eyn, 0.f);
            vset(exyp, 0.f); vset(ezyp, 0.f); vset(hx, 0.f); vset(hy, 0.f); vset(hz, 0.f);
            vset(hxzm, 0.f); vset(hyzm, 0.f);
            vset(ca, ca_u); vset(cb, cb_u);
            float ey_x = 0.f, ez_x = 0.f;
            if (have) {
                lds_ld(ex, lds + L::ex(j));
                lds_ld(ey, lds + L::ey(j));
            }
            if (zin && yin && have_up) lds_ld(exyp, lds + L::ex(j + 1));
            if (own && z > zs) {                       // H(t+1/2) at z - 1, from the plane below
                lds_ld(hxzm, lds + L::hx(j));
                lds_ld(hyzm, lds + L::hy(j));
            }
            if (act) {
                if (!have) {
                    vld(ex, at(fo[0], o));
                    vld(ey, at(fo[1], o));
                }
                if (j == j0) vld(ezc, at(fo[2], o));
                vld(hx, at(fo[3], o));
                vld(hy, at(fo[4], o));
                vld(hz, at(fo[5], o));
                // ez of the row above: this row's H and, passed on, that
                // row's ez (also on the top plane, where E is only copied)
                if (yin) vld(ezyp, at(fo[2], o + sy));
                if (zin) {
                    vld(exn, at(fo[0], o + sz));
                    vld(eyn, at(fo[1], o + sz));
                    if (yin && !have_up) vld(exyp, at(fo[0], o + sy));
                    if (next_mem) {
                        ey_x = *at(fo[1], o + V);
                        ez_x = *at(fo[2], o + V);
                    }
                    if (own && z < uni_z) {
                        vld_stream(ca, at(Ca, o));
                        vld_stream(cb, at(Cb, o));
                    }
                }
            }
            // lane 0: the inputs of H(t+1/2) at x0 - 1 (only its hy and hz
            // are used, so hx's inputs are left out)
            EAround eh;
            float hy_h = 0.f, hz_h = 0.f;
            const bool do_halo = own && halo_x;
            if (do_halo) {
                const uint32_t oh = o - 1;
                eh.ex_c = *at(fo[0], oh);
                eh.ey_c = *at(fo[1], oh);
                eh.ez_c = *at(fo[2], oh);
                eh.ex_n = zin ? *at(fo[0], oh + sz) : 0.f;
                eh.ex_yp = (zin && yin) ? *at(fo[0], oh + sy) : 0.f;
                hy_h = *at(fo[4], oh);
                hz_h = *at(fo[5], oh);
            }
            // ---- H(t+1/2) of the row
            float ey_next = __shfl_down(ey[0], 1, kWave);
            float ez_next = __shfl_down(ezc[0], 1, kWave);
            if (lane == kWave - 1) {
                ey_next = ey_x;
                ez_next = ez_x;
            }
            gfloat *hw_psi = own ? psi_n : nullptr;
#pragma unroll
            for (int c = 0; c < V; ++c) {
                if (xb + c < k.nx) {
                    EAround e;
                    e.ex_c = ex[c]; e.ex_n = exn[c]; e.ex_yp = exyp[c];
                    e.ey_c = ey[c]; e.ey_n = eyn[c];
                    e.ey_xp = c + 1 < V ? ey[c + 1 < V ? c + 1 : 0] : ey_next;
                    e.ez_c = ezc[c]; e.ez_yp = ezyp[c];
                    e.ez_xp = c + 1 < V ? ezc[c + 1 < V ? c + 1 : 0] : ez_next;
                    h_cell(d, k, xb + c, y, z, e, hx[c], hy[c], hz[c], psi_o, hw_psi);
                }
            }
            if (own) {
                // ---- E(t+1) of the row
                float hz_prev = __shfl_up(hz[V - 1], 1, kWave);
                float hy_prev = __shfl_up(hy[V - 1], 1, kWave);
                if (lane == 0) {
                    hz_prev = 0.f;
                    hy_prev = 0.f;
                    if (do_halo) {
                        eh.ey_xp = ey[0];
                        eh.ez_xp = ezc[0];
                        eh.ey_n = eh.ey_c;
                        eh.ez_yp = eh.ez_c;
                        float hx_h = 0.f;
                        h_cell(d, k, x0 - 1, y, z, eh, hx_h, hy_h, hz_h, psi_o, nullptr);
                        hz_prev = hz_h;
                        hy_prev = hy_h;
                    }
                }
                float exu[V], eyu[V], ezu[V];
                vcopy(exu, ex);
                vcopy(eyu, ey);
                vcopy(ezu, ezc);
#pragma unroll
                for (int q = 0; q < V; ++q) {
                    if (xb + q < k.nx) {
                        HAround h;
                        h.hx_c = hx[q]; h.hy_c = hy[q]; h.hz_c = hz[q];
                        h.hx_zm = hxzm[q]; h.hy_zm = hyzm[q];
                        h.hz_ym = hzb[q]; h.hx_ym = hxb[q];
                        h.hz_xm = q > 0 ? hz[q > 0 ? q - 1 : 0] : hz_prev;
                        h.hy_xm = q > 0 ? hy[q > 0 ? q - 1 : 0] : hy_prev;
                        e_cell(d, k, xb + q, y, z, ca[q], cb[q], h, exu[q], eyu[q], ezu[q]);
                    }
                }
                if (act) {
                    vst(at(fn[3], o), hx);
                    vst(at(fn[4], o), hy);
                    vst(at(fn[5], o), hz);
                    vst(at(fn[0], o), exu);
                    vst(at(fn[1], o), eyu);
                    vst(at(fn[2], o), ezu);
                }
            }
            // hand over to the plane above and to the next row
            lds_st(lds + L::ex(j), exn);
            lds_st(lds + L::ey(j), eyn);
            if (j >= 1) {
                lds_st(lds + L::hx(j), hx);
                lds_st(lds + L::hy(j), hy);
            }
            vcopy(hzb, hz);
            vcopy(hxb, hx);
            vcopy(ezc, ezyp);
        }
    }
}

// ---------------------------------------------------------------------------
// Fused stepper, row-parallel form (stepper 2).  A block of R + 1
// wavefronts owns R rows y0 .. y0 + R - 1 of 64 V cells over a z range;
// wavefront w holds row y0 - 1 + w, so wavefront 0 recomputes the halo row
// y0 - 1 (its H(t+1/2) is never stored).  Every wavefront marches z with its
// row's E(t) at z+1 and H(t+1/2) at z in registers, as the split kernels
// carry their planes; H(t+1/2) of the row below reaches the E update through
// LDS, double-buffered by plane parity, with one barrier per plane.  Lane 0
// recomputes H(t+1/2) at x0 - 1.  Waves whose row lies outside the grid keep
// to the barriers and touch no memory.
// ---------------------------------------------------------------------------
template <int V, int R>
__global__ void __launch_bounds__(64 * (R + 1))
k_step_rows(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles, int shape) {
    __shared__ float lds_h[2][R][2][kWave * V];        // [plane parity][row 0..R-1][hx, hz][cells]
    const int4 tl = tiles[block_tile(shape, ntiles)];
    const Desc *__restrict__ d = descs + (tl.x & 0xffff);   // low 16 bits: member index
    const int t = load_step(d->step);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    note_step(d, tl, w, t);
    if (t >= d->nt) return;                            // uniform per block
    const StepK k = step_k(d, t);
    const int lane = threadIdx.x;
    const int x0 = tl.y, y0 = tl.z;
    const int y = y0 - 1 + w;
    const bool row = y >= 0 && y < k.ny;               // uniform per wavefront
    const bool own = w >= 1 && row;
    const int xb = x0 + V * lane;
    const bool act = row && xb < k.nx;
    const bool next_mem = row && lane == kWave - 1 && xb + V < k.nx;
    const bool halo_x = own && lane == 0 && x0 > 0;     // recompute H(t+1/2) at x0 - 1
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const int zs = z0 > 0 ? z0 - 1 : 0;                // plane z0 - 1: H only, not stored
    const int p = t & 1;
    const uint32_t sy = k.sy, sz = k.sz;
    const gfloat *fo[6];
    gfloat *fn[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        fo[c] = gp(d->f[p][c]);
        fn[c] = gp(d->f[1 - p][c]);
    }
    const gfloat *psi_o = gp(d->psi[p][pHxz_d]);
    gfloat *psi_n = gp(d->psi[1 - p][pHxz_d]);
    const gfloat *Ca = gp(d->ca), *Cb = gp(d->cb);
    const int uni_z = d->uni_z;
    const float ca_u = d->ca_u, cb_u = d->cb_u;
    const bool yin = y < k.ny - 1;
    const bool ylo = y >= 1;

    uint32_t o = (uint32_t)zs * sz + (uint32_t)(row ? y : 0) * sy + (uint32_t)xb;
    float exc[V], eyc[V], hxzm[V], hyzm[V];            // E(t) at z; H(t+1/2) at z - 1
    vset(exc, 0.f); vset(eyc, 0.f); vset(hxzm, 0.f); vset(hyzm, 0.f);
    if (act) {
        vld(exc, at(fo[0], o));
        vld(eyc, at(fo[1], o));
    }
#pragma unroll 1
    for (int z = zs; z < z1; ++z, o += sz) {
        const bool zin = z < k.nz - 1;
        const bool hplane = z < z0;
        const bool upd = own && !hplane;               // this row's cells are updated and stored
        // ---- loads: H half of the row, then what the E half adds
        HPlane<V> P;
        h_plane_load(P, fo[0], fo[1], fo[2], fo[3], fo[4], fo[5], o, sy, sz, act, zin, yin, next_mem);
        float ezt[V], ca[V], cb[V];
        vcopy(ezt, P.ez);
        vset(ca, ca_u);
        vset(cb, cb_u);
        if (upd && act) {
            if (!zin) vld(ezt, at(fo[2], o));          // top plane: E is copied
            if (zin && z < uni_z) {
                vld_stream(ca, at(Ca, o));
                vld_stream(cb, at(Cb, o));
            }
        }
        EAround eh;
        float hy_h = 0.f, hz_h = 0.f;
        if (halo_x && !hplane) {
            const uint32_t oh = o - 1;
            eh.ex_c = *at(fo[0], oh);
            eh.ey_c = *at(fo[1], oh);
            eh.ez_c = *at(fo[2], oh);
            eh.ex_n = zin ? *at(fo[0], oh + sz) : 0.f;
            eh.ex_yp = (zin && yin) ? *at(fo[0], oh + sy) : 0.f;
            hy_h = *at(fo[4], oh);
            hz_h = *at(fo[5], oh);
        }
        // ---- H(t+1/2) of the row
        float ey_next = __shfl_down(eyc[0], 1, kWave);
        float ez_next = __shfl_down(P.ez[0], 1, kWave);
        if (lane == kWave - 1) {
            ey_next = P.ey_x;
            ez_next = P.ez_x;
        }
        gfloat *hw_psi = upd ? psi_n : nullptr;
        if (row) {
#pragma unroll
            for (int c = 0; c < V; ++c) {
                if (xb + c < k.nx) {
                    EAround e;
                    e.ex_c = exc[c]; e.ex_n = P.exn[c]; e.ex_yp = P.exyp[c];
                    e.ey_c = eyc[c]; e.ey_n = P.eyn[c];
                    e.ey_xp = c + 1 < V ? eyc[c + 1 < V ? c + 1 : 0] : ey_next;
                    e.ez_c = P.ez[c]; e.ez_yp = P.ezyp[c];
                    e.ez_xp = c + 1 < V ? P.ez[c + 1 < V ? c + 1 : 0] : ez_next;
                    h_cell(d, k, xb + c, y, z, e, P.hx[c], P.hy[c], P.hz[c], psi_o, hw_psi);
                }
            }
        }
        const int par = z & 1;
        if (w < R) {                                   // the row above reads hx, hz of this row
            lds_st(&lds_h[par][w][0][V * lane], P.hx);
            lds_st(&lds_h[par][w][1][V * lane], P.hz);
        }
        __syncthreads();
        if (upd) {
            // ---- E(t+1) of the row
            float hxym[V], hzym[V];
            vset(hxym, 0.f);
            vset(hzym, 0.f);
            if (ylo) {
                lds_ld(hxym, &lds_h[par][w - 1][0][V * lane]);
                lds_ld(hzym, &lds_h[par][w - 1][1][V * lane]);
            }
            float hz_prev = __shfl_up(P.hz[V - 1], 1, kWave);
            float hy_prev = __shfl_up(P.hy[V - 1], 1, kWave);
            if (lane == 0) {
                hz_prev = 0.f;
                hy_prev = 0.f;
                if (halo_x) {
                    eh.ey_xp = eyc[0];
                    eh.ez_xp = P.ez[0];
                    eh.ey_n = eh.ey_c;
                    eh.ez_yp = eh.ez_c;
                    float hx_h = 0.f;
                    h_cell(d, k, x0 - 1, y, z, eh, hx_h, hy_h, hz_h, psi_o, nullptr);
                    hz_prev = hz_h;
                    hy_prev = hy_h;
                }
            }
            float exu[V], eyu[V], ezu[V];
            vcopy(exu, exc);
            vcopy(eyu, eyc);
            vcopy(ezu, ezt);
#pragma unroll
            for (int q = 0; q < V; ++q) {
                if (xb + q < k.nx) {
                    HAround h;
                    h.hx_c = P.hx[q]; h.hy_c = P.hy[q]; h.hz_c = P.hz[q];
                    h.hx_zm = hxzm[q]; h.hy_zm = hyzm[q];
                    h.hz_ym = hzym[q]; h.hx_ym = hxym[q];
                    h.hz_xm = q > 0 ? P.hz[q > 0 ? q - 1 : 0] : hz_prev;
                    h.hy_xm = q > 0 ? P.hy[q > 0 ? q - 1 : 0] : hy_prev;
                    e_cell(d, k, xb + q, y, z, ca[q], cb[q], h, exu[q], eyu[q], ezu[q]);
                }
            }
            if (act) {
                vst(at(fn[3], o), P.hx);
                vst(at(fn[4], o), P.hy);
                vst(at(fn[5], o), P.hz);
                vst(at(fn[0], o), exu);
                vst(at(fn[1], o), eyu);
                vst(at(fn[2], o), ezu);
            }
        }
        // hand over to the plane above
        vcopy(exc, P.exn);
        vcopy(eyc, P.eyn);
        vcopy(hxzm, P.hx);
        vcopy(hyzm, P.hy);
    }
}

template __global__ void k_step_rows<1, 3>(const Desc *, const int4 *, int, int);
template __global__ void k_step_rows<1, 7>(const Desc *, const int4 *, int, int);
template __global__ void k_step_rows<2, 3>(const Desc *, const int4 *, int, int);
template __global__ void k_step_rows<2, 7>(const Desc *, const int4 *, int, int);

template __global__ void k_step_strip<1, 2>(const Desc *, const int4 *, int, int);
template __global__ void k_step_strip<1, 4>(const Desc *, const int4 *, int, int);
template __global__ void k_step_strip<1, 8>(const Desc *, const int4 *, int, int);
template __global__ void k_step_strip<2, 2>(const Desc *, const int4 *, int, int);
template __global__ void k_step_strip<2, 4>(const Desc *, const int4 *, int, int);
template __global__ void k_step_strip<4, 2>(const Desc *, const int4 *, int, int);

// ---------------------------------------------------------------------------
// Fused stepper, halo form (stepper 3).  The split kernels' blocks and
// tiles (xw wavefronts side by side along x on each of ty rows, one row per
// wavefront over a z chunk), with H(t+1/2) and E(t+1) in one kernel.  Rows
// are independent: each wavefront recomputes from the old level what it
// needs of its neighbours' H(t+1/2), bit-identical to the owner's value and
// never stored, so there is no LDS hand-over and no barrier:
//   - hx, hz of the row below (y - 1), the E update's y - 1 operands, on
//     every plane;
//   - hx, hy of the row at z0 - 1, the chunk's first z - 1 operands;
//   - hy, hz at x0 - 1, by lane 0.
// Per plane a lane loads E(t) at z + 1 and the old H of its row and of the
// row below, ez and ex of the row above, and Ca, Cb; E(t) at z of both rows
// and H(t+1/2) at z - 1 of its row stay in registers from the plane below.
// The neighbour rows' values are the other wavefronts' own loads and mostly
// hit L2.  Fields and H-CPML state are double-buffered (step t reads set
// t&1, writes set 1-(t&1)); the E-CPML state is updated in place, as only
// its cell's lane touches it.  The wavefronts covering a DFT x-face column
// copy it to Desc::tapx as the split E kernel does.
// ---------------------------------------------------------------------------
template <int V, bool DIRECT>
__global__ void __launch_bounds__(512)
k_step_halo(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles, int shape, TileGrid tg,
            const int *__restrict__ stepp) {
    int4 tl;
    if (!split_tile<DIRECT>(tiles, ntiles, shape, tg, tl)) return;
    const Desc *__restrict__ d = descs + (tl.x & 0xffff);
    // the step selects the buffer sets: kept uniform (scalar), so the field
    // pointers stay in SGPRs
    const int t = DIRECT ? *stepp : __builtin_amdgcn_readfirstlane(load_step(d->step));
    const StepK kg = step_geom(d);
    const StepTime st = step_time(d);
    scalars_resident(kg, st);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    note_step(d, tl, w, t);
    const int lxw = DIRECT ? shape & 3 : tl.x >> 16;   // listed tiles carry their block width
    const int y = tl.z + (w >> lxw);
    const int nx = kg.nx, ny = kg.ny, nz = kg.nz;
    const int x0 = tl.y + V * ((w & ((1 << lxw) - 1)) << 6);   // the wavefront's first cell
    if (y >= ny || x0 >= nx || t >= st.nt) return;    // uniform per wavefront
    const StepK k = step_at(kg, st, t);
    const int lane = threadIdx.x;
    const int xb = x0 + V * lane;                      // the lane's first cell
    const bool act = xb < nx;
    const bool next_mem = lane == kWave - 1 && xb + V < nx;
    const bool halo_x = lane == 0 && x0 > 0;
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const uint32_t sy = k.sy, sz = k.sz;
    const int p = t & 1;
    const gfloat *fo[6];
    gfloat *fn[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        fo[c] = gp(d->f[p][c]);
        fn[c] = gp(d->f[1 - p][c]);
    }
    const gfloat *psi_o = gp(d->psi[p][pHxz_d]);      // H-CPML state of this step
    gfloat *psi_n = gp(d->psi[1 - p][pHxz_d]);        // and of the next
    const gfloat *Ca = gp(d->ca), *Cb = gp(d->cb);
    const int uni_z = d->uni_z;
    const float ca_u = d->ca_u, cb_u = d->cb_u;
    const bool ylo = y >= 1, yin = y < ny - 1;
    gchar *const tap = d->tapx && wave_taps<V>(x0, nx, k.np) ? (gchar *)d->tapx : nullptr;

    uint32_t o = (uint32_t)z0 * sz + (uint32_t)y * sy + (uint32_t)xb;
    // E(t) at z of the row and of the row below; H(t+1/2) at z - 1 of the row
    float exc[V], eyc[V], exm[V], eym[V], hxzm[V], hyzm[V];
    vset(exc, 0.f); vset(eyc, 0.f); vset(exm, 0.f); vset(eym, 0.f); vset(hxzm, 0.f); vset(hyzm, 0.f);
    if (act) {
        vld(exc, at(fo[0], o));
        vld(eyc, at(fo[1], o));
        if (ylo) {
            vld(exm, at(fo[0], o - sy));
            vld(eym, at(fo[1], o - sy));
        }
    }
    if (z0 > 0) {
        // hx, hy of the row at z0 - 1 (below the top plane: z0 <= nz - 1)
        const uint32_t oh = o - sz;
        float ex[V], ey[V], ez[V], ezyp[V], exyp[V], hz[V];
        vset(ex, 0.f); vset(ey, 0.f); vset(ez, 0.f); vset(ezyp, 0.f); vset(exyp, 0.f); vset(hz, 0.f);
        float ey_x = 0.f, ez_x = 0.f;
        if (act) {
            vld(ex, at(fo[0], oh));
            vld(ey, at(fo[1], oh));
            vld(ez, at(fo[2], oh));
            vld(hxzm, at(fo[3], oh));
            vld(hyzm, at(fo[4], oh));
            if (yin) {
                vld(ezyp, at(fo[2], oh + sy));
                vld(exyp, at(fo[0], oh + sy));
            }
            if (next_mem) {
                ey_x = *at(fo[1], oh + V);
                ez_x = *at(fo[2], oh + V);
            }
        }
        float ey_next = lane_next(ey[0]);
        float ez_next = lane_next(ez[0]);
        if (lane == kWave - 1) {
            ey_next = ey_x;
            ez_next = ez_x;
        }
#pragma unroll
        for (int c = 0; c < V; ++c) {
            if (xb + c < nx) {
                EAround e;
                e.ex_c = ex[c]; e.ex_n = exc[c]; e.ex_yp = exyp[c];
                e.ey_c = ey[c]; e.ey_n = eyc[c];
                e.ey_xp = c + 1 < V ? ey[c + 1 < V ? c + 1 : 0] : ey_next;
                e.ez_c = ez[c]; e.ez_yp = ezyp[c];
                e.ez_xp = c + 1 < V ? ez[c + 1 < V ? c + 1 : 0] : ez_next;
                h_cell(d, k, xb + c, y, z0 - 1, e, hxzm[c], hyzm[c], hz[c], psi_o, nullptr);
            }
        }
    }
    // One plane of the march.  IN (a std::integral_constant): every cell
    // the wavefront updates or recomputes on this plane lies inside the
    // absorber's inner boundary (the wavefront's cells, its row and the row
    // below, the cell left of it), so all lanes are active, every
    // neighbour exists, and the cell updates reduce to the Yee step and the
    // injection (h_cell / e_cell INNER): the same arithmetic, without the
    // per-cell boundary and absorber tests.
    auto plane = [&](int z, auto IN) {
        constexpr bool I = decltype(IN)::value;
        const bool zin = I || z < nz - 1;
        const bool lact = I || act, lyin = I || yin, lylo = I || ylo;
        // ---- loads: the row (H old, E(t) at z + 1, ez at z), the row above
        // (ez, ex), the row below (hx, hz old, E(t) at z + 1, ez), Ca, Cb.
        // Values a lane does not load are never read for a stored result
        // (inactive lanes, missing neighbours), but are zeroed as the split
        // kernels do, except on interior planes, where every lane loads all.
        float exn[V], eyn[V], ez[V], ezyp[V], exyp[V], hx[V], hy[V], hz[V], ca[V], cb[V];
        float exmn[V], eymn[V], ezm[V], hxm[V], hzm[V];
        if (!I) {
            vset(exn, 0.f); vset(eyn, 0.f); vset(ez, 0.f); vset(ezyp, 0.f); vset(exyp, 0.f);
            vset(hx, 0.f); vset(hy, 0.f); vset(hz, 0.f);
            vset(exmn, 0.f); vset(eymn, 0.f); vset(ezm, 0.f); vset(hxm, 0.f); vset(hzm, 0.f);
        }
        vset(ca, ca_u);
        vset(cb, cb_u);
        float ey_x = 0.f, ez_x = 0.f, eym_x = 0.f;
        if (lact) {
            vld(hx, at(fo[3], o));
            vld(hy, at(fo[4], o));
            vld(hz, at(fo[5], o));
            vld(ez, at(fo[2], o));
            if (zin) {
                vld(exn, at(fo[0], o + sz));
                vld(eyn, at(fo[1], o + sz));
                if (lyin) {
                    vld(ezyp, at(fo[2], o + sy));
                    vld(exyp, at(fo[0], o + sy));
                }
                if (next_mem) {
                    ey_x = *at(fo[1], o + V);
                    ez_x = *at(fo[2], o + V);
                }
                if (lylo) {
                    const uint32_t om = o - sy;
                    vld(hxm, at(fo[3], om));
                    vld(hzm, at(fo[5], om));
                    vld(exmn, at(fo[0], om + sz));
                    vld(eymn, at(fo[1], om + sz));
                    vld(ezm, at(fo[2], om));
                    if (next_mem) eym_x = *at(fo[1], om + V);
                }
                if (z < uni_z) {
                    vld_stream(ca, at(Ca, o));
                    vld_stream(cb, at(Cb, o));
                }
            }
        }
        // lane 0: the inputs of hy, hz at x0 - 1
        EAround eh = {};
        float hy_h = 0.f, hz_h = 0.f;
        const bool do_halo = halo_x && zin;
        if (do_halo) {
            const uint32_t oh = o - 1;
            eh.ex_c = *at(fo[0], oh);
            eh.ey_c = *at(fo[1], oh);
            eh.ez_c = *at(fo[2], oh);
            eh.ex_n = *at(fo[0], oh + sz);
            eh.ex_yp = lyin ? *at(fo[0], oh + sy) : 0.f;
            hy_h = *at(fo[4], oh);
            hz_h = *at(fo[5], oh);
        }
        // ---- H(t+1/2) of the row, then hx, hz of the row below
        float ey_next = lane_next(eyc[0]);
        float ez_next = lane_next(ez[0]);
        float eym_next = lane_next(eym[0]);
        if (lane == kWave - 1) {
            ey_next = ey_x;
            ez_next = ez_x;
            eym_next = eym_x;
        }
#pragma unroll
        for (int c = 0; c < V; ++c) {
            if (I || xb + c < nx) {
                EAround e;
                e.ex_c = exc[c]; e.ex_n = exn[c]; e.ex_yp = exyp[c];
                e.ey_c = eyc[c]; e.ey_n = eyn[c];
                e.ey_xp = c + 1 < V ? eyc[c + 1 < V ? c + 1 : 0] : ey_next;
                e.ez_c = ez[c]; e.ez_yp = ezyp[c];
                e.ez_xp = c + 1 < V ? ez[c + 1 < V ? c + 1 : 0] : ez_next;
                h_cell<I>(d, k, xb + c, y, z, e, hx[c], hy[c], hz[c], psi_o, psi_n);
            }
        }
        if (lylo && zin) {
#pragma unroll
            for (int c = 0; c < V; ++c) {
                if (I || xb + c < nx) {
                    // hy of the row below is not needed: its operands
                    // (ex at z + 1, ez at x + 1) are left out
                    EAround e;
                    e.ex_c = exm[c]; e.ex_n = exm[c]; e.ex_yp = exc[c];
                    e.ey_c = eym[c]; e.ey_n = eymn[c];
                    e.ey_xp = c + 1 < V ? eym[c + 1 < V ? c + 1 : 0] : eym_next;
                    e.ez_c = ezm[c]; e.ez_yp = ez[c]; e.ez_xp = ezm[c];
                    float hy_m = 0.f;
                    h_cell<I>(d, k, xb + c, y - 1, z, e, hxm[c], hy_m, hzm[c], psi_o, nullptr);
                }
            }
        }
        // ---- E(t+1) of the row
        float hz_prev = lane_prev(hz[V - 1]);
        float hy_prev = lane_prev(hy[V - 1]);
        if (lane == 0) {
            hz_prev = 0.f;
            hy_prev = 0.f;
            if (do_halo) {
                // only hy, hz of this cell are used: hx's operands are left out
                eh.ey_xp = eyc[0];
                eh.ez_xp = ez[0];
                eh.ey_n = eh.ey_c;
                eh.ez_yp = eh.ez_c;
                float hx_h = 0.f;
                h_cell<I>(d, k, x0 - 1, y, z, eh, hx_h, hy_h, hz_h, psi_o, nullptr);
                hz_prev = hz_h;
                hy_prev = hy_h;
            }
        }
        float exu[V], eyu[V], ezu[V];
        vcopy(exu, exc);
        vcopy(eyu, eyc);
        vcopy(ezu, ez);
#pragma unroll
        for (int q = 0; q < V; ++q) {
            if (I || xb + q < nx) {
                HAround h;
                h.hx_c = hx[q]; h.hy_c = hy[q]; h.hz_c = hz[q];
                h.hx_zm = hxzm[q]; h.hy_zm = hyzm[q];
                h.hz_ym = hzm[q]; h.hx_ym = hxm[q];
                h.hz_xm = q > 0 ? hz[q > 0 ? q - 1 : 0] : hz_prev;
                h.hy_xm = q > 0 ? hy[q > 0 ? q - 1 : 0] : hy_prev;
                e_cell<I>(d, k, xb + q, y, z, ca[q], cb[q], h, exu[q], eyu[q], ezu[q]);
            }
        }
        if (lact) {
            vst_stream(at(fn[3], o), hx);
            vst_stream(at(fn[4], o), hy);
            vst_stream(at(fn[5], o), hz);
            vst_stream(at(fn[0], o), exu);
            vst_stream(at(fn[1], o), eyu);
            vst_stream(at(fn[2], o), ezu);
        }
        if (!I && tap) e_tap<V>(tap, xb, nx, k.np, nz, ny, z, y, eyu, ezu, hy, hz, hy_prev, hz_prev);
        // hand over to the plane above
        vcopy(exc, exn);
        vcopy(eyc, eyn);
        vcopy(exm, exmn);
        vcopy(eym, eymn);
        vcopy(hxzm, hx);
        vcopy(hyzm, hy);
    };
    // interior wavefront: its cells and the cell left of it lie in
    // [np, nx - np) along x, its row and the row below in [np, ny - np),
    // and it copies no DFT column; its planes np .. nz - np - 1 are interior
    const int np = k.np;
    const bool wave_in = !tap && x0 >= np && x0 + kWave * V <= nx - np && y >= np && y < ny - np;
#pragma unroll 1
    for (int z = z0; z < z1; ++z, o += sz) {
        if (wave_in && z >= np && z < nz - np)
            plane(z, std::true_type{});
        else
            plane(z, std::false_type{});
    }
}

template __global__ void k_step_halo<1, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_step_halo<1, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_step_halo<2, false>(const Desc *, const int4 *, int, int, TileGrid, const int *);
template __global__ void k_step_halo<2, true>(const Desc *, const int4 *, int, int, TileGrid, const int *);

// ---------------------------------------------------------------------------
// Running single-frequency DFT on the phasor planes
// (cpu/field_updates.py:373-491, called at cpu/main_sim.py:166-195 and
// cpu/incident_sim.py:159-161).  E samples are 2-point averages, H samples
// 4-point averages, summed in float32 in the reference's order.
//
// FDTD_DFT_REFERENCE: acc = float(double(acc) + 0.5*double(sum) * cos/sin),
// the CPU path's complex64-store / complex128-add semantics, bit for bit.
// FDTD_DFT_COMPENSATED: float-float running sum; the product sum * w is
// split exactly with an FMA (w itself as a float hi/lo pair, scaled by the
// averaging factor) and added with TwoSum, so the accumulated value carries
// ~44 significant bits in float32 arithmetic only.
//
// The first block then moves the step counter to the next step (note_step).
// ---------------------------------------------------------------------------
// Running sums are read and written once per step, streamed (vld_stream).
__device__ __forceinline__ float4 acc_load(const float4 &v) {
    const vf4 q = __builtin_nontemporal_load(reinterpret_cast<const vf4 *>(&v));
    return make_float4(q.x, q.y, q.z, q.w);
}

__device__ __forceinline__ void acc_store(float4 &v, const float4 &r) {
    __builtin_nontemporal_store(vf4{r.x, r.y, r.z, r.w}, reinterpret_cast<vf4 *>(&v));
}

__device__ __forceinline__ void acc_ref(float4 &a, float s, double half, double c, double sn) {
    const double avg = half * (double)s;
    a.x = (float)((double)a.x + avg * c);
    a.y = (float)((double)a.y + avg * sn);
}

__device__ __forceinline__ void ff_add(float &hi, float &lo, float a, float th, float tl) {
    const float p = a * th;
    float pe = fmaf(a, th, -p);
    pe = fmaf(a, tl, pe);
    const float s = hi + p;
    const float bb = s - hi;
    const float e = (hi - (s - bb)) + (p - bb);
    const float l = lo + (e + pe);
    const float s2 = s + l;
    lo = l - (s2 - s);
    hi = s2;
}

__device__ __forceinline__ void acc_comp(float4 &a, float s, float4 w) {
    ff_add(a.x, a.z, s, w.x, w.y);   // re: cos
    ff_add(a.y, a.w, s, w.z, w.w);   // im: sin
}

__global__ void __launch_bounds__(256)
k_dft(const Desc *__restrict__ descs) {
    const Desc *__restrict__ d = descs + blockIdx.y;
    const int t = load_step(d->step + 1);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0 && t < d->nt) __hip_atomic_store(d->step, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t >= d->nt || i >= d->dft_items) return;
    const int nsamp = (int)d->face_off[d->nfaces];
    const long long sy = d->sy, sz = d->sz;
    // the step's kernels have written set current_set(t + 1)
    const int cs = current_set(*d, t + 1);
    const float *__restrict__ Ex = d->f[cs][0];
    const float *__restrict__ Ey = d->f[cs][1];
    const float *__restrict__ Ez = d->f[cs][2];
    const float *__restrict__ Hx = d->f[cs][3];
    const float *__restrict__ Hy = d->f[cs][4];
    const float *__restrict__ Hz = d->f[cs][5];
    if (i < nsamp) {
        int f = 0;
        while (f + 1 < d->nfaces && i >= (int)d->face_off[f + 1]) ++f;
        const int j = i - (int)d->face_off[f];
        const int nb = d->face_nb[f];
        const int a = j / nb, b = j - a * nb;
        if (a < 1 || b < 1) return;
        const int np = d->npml, pos = d->face_pos[f];
        float s[4];
        switch (d->face_kind[f]) {
        case kFaceZ: {   // (a, b) = (y, x); ex ey hx hy
            const long long o = (long long)pos * sz + (long long)(np + a) * sy + (np + b);
            s[0] = Ex[o] + Ex[o - 1];
            s[1] = Ey[o] + Ey[o - sy];
            s[2] = Hx[o] + Hx[o - sy] + Hx[o - sz] + Hx[o - sz - sy];
            s[3] = Hy[o] + Hy[o - 1] + Hy[o - sz] + Hy[o - sz - 1];
            break;
        }
        case kFaceX: {   // (a, b) = (z, y); ey ez hy hz
            if (d->tapx && (d->fused == 0 || d->fused == 3)) {
                // the split E kernel's (or halo stepper's) copies of the face columns:
                // (ey, ez, hy(x-1), hy), (hz(x-1), hz, -, -) per (z, y)
                const int ny = d->ny;
                const float4 *T = d->tapx + 2 * (pos == np + 1 ? 0 : (long long)d->nz * ny);
                const long long ti = (long long)(np + a) * ny + (np + b);
                const float4 c0 = T[2 * ti], c1 = T[2 * ti + 1];
                const float4 y0 = T[2 * (ti - 1)], y1 = T[2 * (ti - 1) + 1];
                const float4 z0 = T[2 * (ti - ny)];
                s[0] = c0.x + y0.x;
                s[1] = c0.y + z0.y;
                s[2] = c0.w + c0.z + z0.w + z0.z;
                s[3] = c1.y + c1.x + y1.y + y1.x;
                break;
            }
            const long long o = (long long)(np + a) * sz + (long long)(np + b) * sy + pos;
            s[0] = Ey[o] + Ey[o - sy];
            s[1] = Ez[o] + Ez[o - sz];
            s[2] = Hy[o] + Hy[o - 1] + Hy[o - sz] + Hy[o - sz - 1];
            s[3] = Hz[o] + Hz[o - 1] + Hz[o - sy] + Hz[o - sy - 1];
            break;
        }
        default: {       // kFaceY, (a, b) = (z, x); ex ez hx hz
            const long long o = (long long)(np + a) * sz + (long long)pos * sy + (np + b);
            s[0] = Ex[o] + Ex[o - 1];
            s[1] = Ez[o] + Ez[o - sz];
            s[2] = Hx[o] + Hx[o - sy] + Hx[o - sz] + Hx[o - sz - sy];
            s[3] = Hz[o] + Hz[o - sy] + Hz[o - 1] + Hz[o - sy - 1];
            break;
        }
        }
        float4 *acc = d->acc;
        if (d->dft_accum == FDTD_DFT_REFERENCE) {
            const double4 w = d->tw64[t];
            for (int c = 0; c < 4; ++c) {
                float4 &v = acc[d->acc_off[4 * f + c] + j];
                float4 r = acc_load(v);
                if (c < 2) acc_ref(r, s[c], 0.5, w.x, w.y);
                else acc_ref(r, s[c], 0.25, w.z, w.w);
                acc_store(v, r);
            }
        } else {
            const float4 we = d->tw32[2 * t], wh = d->tw32[2 * t + 1];
            for (int c = 0; c < 4; ++c) {
                float4 &v = acc[d->acc_off[4 * f + c] + j];
                float4 r = acc_load(v);
                acc_comp(r, s[c], c < 2 ? we : wh);
                acc_store(v, r);
            }
        }
    } else if (d->rec) {
        // INCIDENT: record the read-out plane of ex, ey, hx, hy at step t
        // (cpu/incident_sim.py:153-156).
        const int k = i - nsamp;
        const int plane = d->ny * d->nx;
        const int c = k / plane;
        const int r = k - c * plane;
        const int yy = r / d->nx, xx = r - yy * d->nx;
        const int src_comp[4] = {0, 1, 3, 4};
        const float v = d->f[cs][src_comp[c]][(long long)d->read_out * sz + (long long)yy * sy + xx];
        d->rec[c * d->beam_cs + (long long)t * plane + r] = v;
    }
}

// Phasor read-out: complex64 value of each sample, float(hi + lo).
__global__ void __launch_bounds__(256)
k_phasor_out(const float4 *__restrict__ acc, long long n, float2 *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 v = acc[i];
        out[i] = make_float2((float)((double)v.x + (double)v.z), (float)((double)v.y + (double)v.w));
    }
}

// ---------------------------------------------------------------------------
// Streaming copy probe: one element of width V floats per thread, the same
// coalesced access the stencil kernels make with V = 1.  Measures the
// practical copy bandwidth and calibrates the PMC byte counters.
// ---------------------------------------------------------------------------
template <int V>
__global__ void __launch_bounds__(256)
k_copy_probe(const float *__restrict__ src, float *__restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (V == 4) {
        if (i < n / 4) reinterpret_cast<float4 *>(dst)[i] = reinterpret_cast<const float4 *>(src)[i];
    } else if (i < n) {
        dst[i] = src[i];
    }
}

template __global__ void k_copy_probe<1>(const float *, float *, long long);
template __global__ void k_copy_probe<4>(const float *, float *, long long);

// Multi-stream probe: each thread reads NR arrays and writes NW arrays of n
// floats at its own index, 4 bytes per lane per stream -- the HBM access mix
// of a stencil kernel (H half: 6 reads, 3 writes; E half: 8 and 3; fused
// step: 8 and 6) without its neighbour reuse or arithmetic.  Its bandwidth
// is the practical ceiling for that mix.
template <int NR, int NW>
__global__ void __launch_bounds__(256)
k_stream_probe(const float *__restrict__ in, float *__restrict__ out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.f;
#pragma unroll
    for (int s = 0; s < NR; ++s) acc += in[s * n + i];
#pragma unroll
    for (int s = 0; s < NW; ++s) out[s * n + i] = acc + (float)s;
}

// The split kernels' access pattern without their arithmetic, for the
// practical ceiling of that pattern: the same arrays (field stride apart),
// blocks of 256 lanes along x (1 KB per array), tiles in plain order, a z
// march over zchunk planes with the plane-to-plane carry in registers, and
// the y +- 1 row loads.  KIND 0 is the H half (reads ex ey ez hx hy hz, ex
// and ez of row y+1, ex ey of plane z+1; writes hx hy hz), KIND 1 the E half
// (reads ex ey ez hx hy hz ca cb, hz and hx of row y-1, hx hy of plane z-1;
// writes ex ey ez).  KIND 2 is the halo stepper's one kernel per step
// (reads ex ey ez hx hy hz ca cb of its row, ex ey ez hx hz of row y-1, ex ez
// of row y+1; writes the six fields to the other buffer set, streamed).
// Algorithmic streams: 9, 11 and 14 floats per cell.
template <int KIND>
__global__ void __launch_bounds__(256)
k_stencil_probe(float *__restrict__ base, long long stride, int nx, int ny, int nz, uint32_t sy, uint32_t sz,
                int zchunk, int ntx) {
    const int b = blockIdx.x;
    const int xt = b % ntx, rest = b / ntx;
    const int y = rest % ny, zc = rest / ny;
    const int x = xt * 256 + (int)threadIdx.x;
    if (x >= nx) return;
    const int z0 = zc * zchunk, z1 = min(nz, z0 + zchunk);
    gfloat *A0 = gp(base), *A1 = gp(base + stride), *A2 = gp(base + 2 * stride);
    gfloat *A3 = gp(base + 3 * stride), *A4 = gp(base + 4 * stride), *A5 = gp(base + 5 * stride);
    uint32_t o = (uint32_t)z0 * sz + (uint32_t)y * sy + (uint32_t)x;
    if (KIND == 0) {
        const bool yin = y + 1 < ny;
        float c0 = *at(A0, o), c1 = *at(A1, o);
        for (int z = z0; z < z1; ++z, o += sz) {
            const bool zin = z + 1 < nz;
            const float n0 = zin ? *at(A0, o + sz) : 0.f, n1 = zin ? *at(A1, o + sz) : 0.f;
            const float e2 = *at(A2, o), h3 = *at(A3, o), h4 = *at(A4, o), h5 = *at(A5, o);
            const float p0 = yin ? *at(A0, o + sy) : 0.f, p2 = yin ? *at(A2, o + sy) : 0.f;
            const float acc = ((c0 - n0) + (c1 - n1)) + ((p0 - c0) + (p2 - e2));
            *at(A3, o) = h3 - acc;
            *at(A4, o) = h4 + acc;
            *at(A5, o) = h5 * acc;
            c0 = n0;
            c1 = n1;
        }
    } else if (KIND == 2) {
        const gfloat *A6 = gp(base + 6 * stride), *A7 = gp(base + 7 * stride);
        gfloat *B[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) B[c] = gp(base + (8 + c) * stride);
        const bool ylo = y >= 1, yin = y + 1 < ny;
        float exc = *at(A0, o), eyc = *at(A1, o);
        float exm = ylo ? *at(A0, o - sy) : 0.f, eym = ylo ? *at(A1, o - sy) : 0.f;
        float hxz = 0.f, hyz = 0.f;
        for (int z = z0; z < z1; ++z, o += sz) {
            const bool zin = z + 1 < nz;
            const float exn = zin ? *at(A0, o + sz) : 0.f, eyn = zin ? *at(A1, o + sz) : 0.f;
            const float ez = *at(A2, o);
            float hx = *at(A3, o), hy = *at(A4, o), hz = *at(A5, o);
            const float ca = __builtin_nontemporal_load(at(A6, o)), cb = __builtin_nontemporal_load(at(A7, o));
            const float ezp = yin ? *at(A2, o + sy) : 0.f, exp_ = yin ? *at(A0, o + sy) : 0.f;
            float hxm = 0.f, hzm = 0.f, exmn = 0.f, eymn = 0.f, ezm = 0.f;
            if (ylo) {
                hxm = *at(A3, o - sy);
                hzm = *at(A5, o - sy);
                ezm = *at(A2, o - sy);
                if (zin) {
                    exmn = *at(A0, o - sy + sz);
                    eymn = *at(A1, o - sy + sz);
                }
            }
            const float eyx = __shfl_down(eyc, 1, kWave), ezx = __shfl_down(ez, 1, kWave);
            const float eymx = __shfl_down(eym, 1, kWave);
            hx -= 0.5f * ((ezp - ez) - (eyn - eyc));
            hy -= 0.5f * ((exn - exc) - (ezx - ez));
            hz -= 0.5f * ((eyx - eyc) - (exp_ - exc));
            hxm -= 0.5f * ((ez - ezm) - (eymn - eym));
            hzm -= 0.5f * ((eymx - eym) - (exc - exm));
            const float hyx = __shfl_up(hy, 1, kWave), hzx = __shfl_up(hz, 1, kWave);
            __builtin_nontemporal_store(hx, at(B[3], o));
            __builtin_nontemporal_store(hy, at(B[4], o));
            __builtin_nontemporal_store(hz, at(B[5], o));
            __builtin_nontemporal_store(ca * exc + cb * ((hz - hzm) - (hy - hyz)), at(B[0], o));
            __builtin_nontemporal_store(ca * eyc + cb * ((hx - hxz) - (hz - hzx)), at(B[1], o));
            __builtin_nontemporal_store(ca * ez + cb * ((hy - hyx) - (hx - hxm)), at(B[2], o));
            exc = exn;
            eyc = eyn;
            exm = exmn;
            eym = eymn;
            hxz = hx;
            hyz = hy;
        }
    } else {
        const gfloat *A6 = gp(base + 6 * stride), *A7 = gp(base + 7 * stride);
        const bool ylo = y >= 1;
        float m3 = z0 >= 1 ? *at(A3, o - sz) : 0.f, m4 = z0 >= 1 ? *at(A4, o - sz) : 0.f;
        for (int z = z0; z < z1; ++z, o += sz) {
            const float h3 = *at(A3, o), h4 = *at(A4, o), h5 = *at(A5, o);
            const float e0 = *at(A0, o), e1 = *at(A1, o), e2 = *at(A2, o);
            const float ca = *at(A6, o), cb = *at(A7, o);
            const float q5 = ylo ? *at(A5, o - sy) : 0.f, q3 = ylo ? *at(A3, o - sy) : 0.f;
            *at(A0, o) = ca * e0 + cb * ((h5 - q5) - (h4 - m4));
            *at(A1, o) = ca * e1 + cb * ((h3 - m3) - h5);
            *at(A2, o) = ca * e2 + cb * ((h4 - h3) - (h3 - q3));
            m3 = h3;
            m4 = h4;
        }
    }
}

template __global__ void k_stencil_probe<0>(float *, long long, int, int, int, uint32_t, uint32_t, int, int);
template __global__ void k_stencil_probe<1>(float *, long long, int, int, int, uint32_t, uint32_t, int, int);
template __global__ void k_stencil_probe<2>(float *, long long, int, int, int, uint32_t, uint32_t, int, int);

template __global__ void k_stream_probe<1, 1>(const float *, float *, long long);
template __global__ void k_stream_probe<6, 3>(const float *, float *, long long);
template __global__ void k_stream_probe<8, 3>(const float *, float *, long long);
template __global__ void k_stream_probe<8, 6>(const float *, float *, long long);

// ---------------------------------------------------------------------------
// Near-to-far transform over the top face (near2far.py:28-156), FP64.
// The integrand is linear in J = (-hy_u, hx_u) and M = (ey_u, -ex_u), so
// each theta needs the four Simpson-weighted sums of J_x, J_y, M_x, M_y
// times exp(jk(x sin th cos ph + y sin th sin ph + z_top cos th)); the phase
// factorises into per-column and per-row tables (k_n2f_tables).
// Block (theta, split) reduces its rows with wavefront shuffles, then LDS.
// ---------------------------------------------------------------------------
struct cd { double re, im; };
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Phase tables of one call: per direction (theta, phi), the Simpson-weighted
// column factors w_i exp(jk x_i sin th cos ph) and row factors
// w_j exp(jk y_j sin th sin ph) with x_i = (i - (nxp-1)/2) dx
// (near2far.py:71-73, :126-131), the top-face factor exp(jk z_top cos th)
// and (cos th, sin th, cos ph, sin ph).  Grid (ndir, ceil(max(nxp, nyp) /
// 256)).
__device__ __forceinline__ double simpson_w(int n, int i) {
    return (i == 0 || i == n - 1) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
}

__global__ void __launch_bounds__(256)
k_n2f_tables(const double *__restrict__ thetas_deg, const double *__restrict__ phis_deg, int nxp, int nyp,
             double dx, double k, double z_top, double2 *__restrict__ phx, double2 *__restrict__ phy,
             double2 *__restrict__ zph, double4 *__restrict__ trig) {
    const int t = blockIdx.x;
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    const double th = thetas_deg[t] * (3.141592653589793 / 180.0);
    const double ph = phis_deg[t] * (3.141592653589793 / 180.0);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    if (i < nxp) {
        const double xv = (double)(i - (nxp - 1) / 2) * dx, w = simpson_w(nxp, i) * dx / 3.0;
        double s, c;
        sincos(k * (xv * (st * cp)), &s, &c);
        phx[(long long)t * nxp + i] = make_double2(w * c, w * s);
    }
    if (i < nyp) {
        const double yv = (double)(i - (nyp - 1) / 2) * dx, w = simpson_w(nyp, i) * dx / 3.0;
        double s, c;
        sincos(k * (yv * (st * sp)), &s, &c);
        phy[(long long)t * nyp + i] = make_double2(w * c, w * s);
    }
    if (i == 0) {
        double s, c;
        sincos(k * (z_top * ct), &s, &c);
        zph[t] = make_double2(c, s);
        trig[t] = make_double4(ct, st, cp, sp);
    }
}

// A top-face sample as complex64: from a staged complex64 plane, or from
// the running-sum accumulators (hi.re, hi.im, lo.re, lo.im) rounded as
// k_phasor_out rounds them, so both sources give the same values.
__device__ __forceinline__ float2 n2f_sample(const float2 *p, long long q) { return p[q]; }
__device__ __forceinline__ float2 n2f_sample(const float4 *p, long long q) {
    const float4 v = p[q];
    return make_float2((float)((double)v.x + (double)v.z), (float)((double)v.y + (double)v.w));
}

template <typename P>
__global__ void __launch_bounds__(256)
k_n2f_partial(const P *__restrict__ ex_u, const P *__restrict__ ey_u,
              const P *__restrict__ hx_u, const P *__restrict__ hy_u,
              int nxp, int nyp, int nsplit,
              const double2 *__restrict__ phx,   // [ntheta][nxp]  w_i exp(jk x_i a)
              const double2 *__restrict__ phy,   // [ntheta][nyp]  w_j exp(jk y_j b)
              double *__restrict__ partial) {    // [ntheta][nsplit][8]
    const int th = blockIdx.x, sp = blockIdx.y;
    const int rows = (nyp + nsplit - 1) / nsplit;
    const int j0 = sp * rows, j1 = min(nyp, j0 + rows);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};     // Jx, Jy, Mx, My (re, im)
    const double2 *px = phx + (long long)th * nxp;
    const double2 *py = phy + (long long)th * nyp;
    for (int j = j0; j < j1; ++j) {
        const cd wy = {py[j].x, py[j].y};
        double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = threadIdx.x; i < nxp; i += blockDim.x) {
            const long long q = (long long)j * nxp + i;
            const cd w = {px[i].x, px[i].y};
            const float2 a = n2f_sample(hy_u, q), b = n2f_sample(hx_u, q);
            const float2 c = n2f_sample(ey_u, q), e = n2f_sample(ex_u, q);
            const cd jx = cmul(w, cd{-(double)a.x, -(double)a.y});
            const cd jy = cmul(w, cd{(double)b.x, (double)b.y});
            const cd mx = cmul(w, cd{(double)c.x, (double)c.y});
            const cd my = cmul(w, cd{-(double)e.x, -(double)e.y});
            r[0] += jx.re; r[1] += jx.im; r[2] += jy.re; r[3] += jy.im;
            r[4] += mx.re; r[5] += mx.im; r[6] += my.re; r[7] += my.im;
        }
        for (int k = 0; k < 8; k += 2) {
            const cd v = cmul(wy, cd{r[k], r[k + 1]});
            acc[k] += v.re;
            acc[k + 1] += v.im;
        }
    }
    __shared__ double red[8][256 / kWave];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    for (int k = 0; k < 8; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) red[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double v = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) v += red[threadIdx.x][w];
        partial[((long long)th * nsplit + sp) * 8 + threadIdx.x] = v;
    }
}

template __global__ void k_n2f_partial<float2>(const float2 *, const float2 *, const float2 *,
                                              const float2 *, int, int, int, const double2 *,
                                              const double2 *, double *);
template __global__ void k_n2f_partial<float4>(const float4 *, const float4 *, const float4 *,
                                              const float4 *, int, int, int, const double2 *,
                                              const double2 *, double *);

// Sum the splits and assemble E/H far fields (near2far.py:94-106).
// trig: per theta (cos th, sin th, cos ph, sin ph); zph: exp(jk z_top cos th);
// wgt: ff_weight = jk exp(-jkr) / (4 pi r).  out: [4][ntheta] complex128.
__global__ void __launch_bounds__(64)
k_n2f_final(const double *__restrict__ partial, int ntheta, int nsplit,
            const double4 *__restrict__ trig, const double2 *__restrict__ zph,
            double2 wgt, double eta0, double2 *__restrict__ out) {
    const int th = blockIdx.x * blockDim.x + threadIdx.x;
    if (th >= ntheta) return;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nsplit; ++k)
        for (int c = 0; c < 8; ++c) s[c] += partial[((long long)th * nsplit + k) * 8 + c];
    const cd ez = {zph[th].x, zph[th].y};
    const cd Jx = cmul(ez, cd{s[0], s[1]}), Jy = cmul(ez, cd{s[2], s[3]});
    const cd Mx = cmul(ez, cd{s[4], s[5]}), My = cmul(ez, cd{s[6], s[7]});
    const double ct = trig[th].x, st_ = trig[th].y, cp = trig[th].z, sp = trig[th].w;
    (void)st_;
    const cd Nt = {Jx.re * ct * cp + Jy.re * ct * sp, Jx.im * ct * cp + Jy.im * ct * sp};
    const cd Np = {-Jx.re * sp + Jy.re * cp, -Jx.im * sp + Jy.im * cp};
    const cd Lt = {Mx.re * ct * cp + My.re * ct * sp, Mx.im * ct * cp + My.im * ct * sp};
    const cd Lp = {-Mx.re * sp + My.re * cp, -Mx.im * sp + My.im * cp};
    const cd w = {wgt.x, wgt.y};
    const cd Et = cmul(w, cd{-Lp.re - eta0 * Nt.re, -Lp.im - eta0 * Nt.im});
    const cd Ep = cmul(w, cd{Lt.re - eta0 * Np.re, Lt.im - eta0 * Np.im});
    const cd Ht = cmul(w, cd{Np.re - Lt.re / eta0, Np.im - Lt.im / eta0});
    const cd Hp = cmul(w, cd{-Nt.re - Lp.re / eta0, -Nt.im - Lp.im / eta0});
    out[0 * ntheta + th] = make_double2(Et.re, Et.im);
    out[1 * ntheta + th] = make_double2(Ep.re, Ep.im);
    out[2 * ntheta + th] = make_double2(Ht.re, Ht.im);
    out[3 * ntheta + th] = make_double2(Hp.re, Hp.im);
}

}  // namespace fdtd
