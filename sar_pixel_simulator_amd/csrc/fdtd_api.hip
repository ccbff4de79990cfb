// C-ABI of libfdtd_hip.so: handles, device memory, launch sequencing and
// HIP graphs around the kernels in fdtd_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fdtd_kernels.h"

namespace fdtd {
template <int V, bool DIRECT, bool FLAT>
__global__ void k_h_half(const Desc *, const int4 *, int, int, TileGrid, const int *);
template <int V, bool DIRECT, bool FLAT>
__global__ void k_e_half(const Desc *, const int4 *, int, int, TileGrid, const int *);
__global__ void k_dft(const Desc *);
template <int V, int R>
__global__ void k_step_strip(const Desc *, const int4 *, int, int);
template <int V, int R>
__global__ void k_step_rows(const Desc *, const int4 *, int, int);
template <int V, bool DIRECT>
__global__ void k_step_halo(const Desc *, const int4 *, int, int, TileGrid, const int *);
__global__ void k_phasor_out(const float4 *, long long, float2 *);
__global__ void k_n2f_tables(const double *, const double *, int, int, double, double, double, double2 *,
                             double2 *, double2 *, double4 *);
template <typename P>
__global__ void k_n2f_partial(const P *, const P *, const P *, const P *, int, int, int, const double2 *,
                              const double2 *, double *);
__global__ void k_n2f_final(const double *, int, int, const double4 *, const double2 *, double2, double,
                            double2 *);
template <int V>
__global__ void k_copy_probe(const float *, float *, long long);
template <int NR, int NW>
__global__ void k_stream_probe(const float *, float *, long long);
template <int KIND>
__global__ void k_stencil_probe(float *, long long, int, int, int, uint32_t, uint32_t, int, int);
__global__ void k_mat_cylinders(const FdtdMatCylinder *, double, int *, long long, long long);
__global__ void k_mat_build(FdtdMatSpec, const int32_t *, const FdtdMatCylinder *, int, int, int, long long,
                            long long, float *, float *);
}  // namespace fdtd

using namespace fdtd;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static thread_local int g_device = 0;

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(FDTD_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

#define CHECK(cond, ...)                                      \
    do {                                                      \
        if (!(cond)) return fail(FDTD_EINVAL, __VA_ARGS__);   \
    } while (0)

// Distance between consecutive field arrays in the field allocation: the
// array size plus 1 MB + 8 KB + 768 B, so that the six (twelve) arrays a
// stencil lane streams from do not start at the same offset modulo large
// powers of two (measured 1.5 % faster than back-to-back arrays).
static long long field_stride(long long ncell) { return ncell + (((1LL << 20) + 8192 + 768) >> 2); }

// Row pitch of the device arrays: nx rounded up to FDTD_PITCH_ALIGN floats
// (environment, default 16 = 64-byte rows; 1 = unpadded, which limits the
// split kernels to one cell per lane).
static long long pitch_of(int nx) {
    static const long long align = [] {
        const char *v = getenv("FDTD_PITCH_ALIGN");
        long long a = v ? atoll(v) : 16;
        return (a >= 1 && a <= 256) ? a : 16;
    }();
    return ((long long)nx + align - 1) / align * align;
}

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct FdtdBeam {
    int device = 0;
    int nt = 0, ny = 0, nx = 0;
    float *d = nullptr;
    std::atomic<int> refs{1};
};

static void beam_unref(FdtdBeam *b) {
    if (b && --b->refs == 0) {
        (void)hipSetDevice(b->device);
        (void)hipFree(b->d);
        delete b;
    }
}

// Launch shape a simulation or batch starts with, per stepper.  Split:
// one x cell per lane, blocks of four wavefronts as wide along x as the
// grid fills, 2-plane chunks, plain block order (measured fastest at
// 761 x 761 x 203).  Halo: one x cell per lane, blocks of 2 x 2
// wavefronts (two rows of 128 cells), 8-plane chunks (fastest there).
// Fused: 4-row strips, 8-plane chunks.
static Launch default_launch(int fused) {
    Launch l;
    if (fused == 3) {
        l.ty = 2;
        l.zchunk = 8;
        l.xvec = 1;
        l.xw = 2;
    } else if (fused == 2) {
        l.ty = 7;
        l.zchunk = 8;
        l.xvec = 1;
        l.xw = 1;
    } else if (fused) {
        l.ty = 4;
        l.zchunk = 8;
        l.xw = 1;
    }
    return l;
}

// A set of simulations stepped together: one launch per kernel per step.
struct Group {
    int device = 0;
    hipStream_t stream = nullptr;
    Desc *d_descs = nullptr;
    int4 *d_tiles = nullptr;
    size_t descs_cap = 0, tiles_cap = 0;   // allocated lengths, reused while they suffice
    int ntiles = 0, nmodels = 0, mode = FDTD_MODE_MAIN, fused = 0;
    std::vector<std::pair<int, int>> slabs;   // split stepper: tile ranges stepped H then E
    int xvec = 1;                  // x cells per lane of the kernels in use
    bool flat = false;             // split stepper on flat tiles (kXwFlat)
    int ty = 1, xw = 1;            // rows and x wavefronts per block (split), rows per strip (fused)
    TileGrid tg = {};              // split kernels, single simulation: tiles from the block index
    const int *step1 = nullptr;    // single simulation: its step counter
    long long max_dft_items = 0;
    Launch launch;
    std::vector<FdtdSim *> members;
    std::map<int, hipGraphExec_t> graphs;
    std::map<int, hipGraph_t> graph_src;
};

struct FdtdSim {
    int device = 0;
    FdtdGeom g{};
    FdtdCoeffs c{};
    Desc h{};                    // host copy of the descriptor
    long long pitch = 0, ncell = 0;
    float *fields = nullptr;     // 6 * ncell per buffer set (MAIN: 2 sets)
    float *mat = nullptr;        // 2 * ncell (Ca, Cb)
    float *psi = nullptr;        // all 24 arrays
    long long psi_len[24] = {};
    float4 *acc = nullptr;
    float4 *tapx = nullptr;      // MAIN: x-face column copies, 2 * 2 * nz * ny
    double4 *tw64 = nullptr;
    float4 *tw32 = nullptr;
    float *src = nullptr;
    int *step = nullptr;
    FdtdBeam *beam = nullptr;
    Group self;                  // batch of one
    void *mat_stage = nullptr;   // device staging of material specs, kept (a free waits for the device)
    size_t mat_stage_len = 0;
    float *vol_stage = nullptr;  // device staging of host volume uploads, kept likewise
    size_t vol_stage_len = 0;
    float2 *scratch_c64 = nullptr;
    long long scratch_len = 0;
    bool mat_set = false;        // materials uploaded or uniform value given
};

// A batch steps its members in one of three ways (fdtd_batch_set_mode):
// packed, every member in one tile list (one launch per kernel per step for
// all of them); serial, each member's own direct launches one member after
// another; hybrid, members below a cell count packed and the rest serial.
// Everything runs on the batch's stream (grp.stream), captured per chunk of
// steps into one graph.
struct FdtdBatch {
    Group grp;                        // the packed members (may be none)
    std::vector<FdtdSim *> all;       // every member, in creation order
    std::vector<FdtdSim *> serial;    // members stepped with their own launches
    std::vector<int> lane;            // serial member k runs on streams[lane[k]]
    std::vector<hipStream_t> streams; // [0] = grp.stream; more for concurrent serial lanes
    std::vector<hipEvent_t> fork;     // capture: per lane, joined back into grp.stream
    int mode = FDTD_BATCH_PACKED;
    int nlanes = 1;
    long long small_cells = 0;
    std::map<int, hipGraphExec_t> graphs;
    std::map<int, hipGraph_t> graph_src;
};

// ---------------------------------------------------------------------------
// geometry bookkeeping
// ---------------------------------------------------------------------------
static int check_geom(const FdtdGeom *g) {
    CHECK(g, "geom is NULL");
    CHECK(g->nx >= 3 && g->ny >= 3 && g->nz >= 3, "grid too small: %d x %d x %d", g->nx, g->ny, g->nz);
    CHECK(g->nz < 65536, "nz %d too large for the tile encoding", g->nz);
    CHECK(g->npml >= 2 && g->npml <= FDTD_MAX_PML, "npml %d outside [2, %d]", g->npml, FDTD_MAX_PML);
    CHECK(2 * g->npml + 2 <= g->nx && 2 * g->npml + 2 <= g->ny && 2 * g->npml + 2 <= g->nz,
          "grid %d x %d x %d too small for npml %d", g->nx, g->ny, g->nz, g->npml);
    CHECK(g->nt >= 0 && g->nt_inc >= 0, "negative step counts");
    CHECK(g->mode == FDTD_MODE_MAIN || g->mode == FDTD_MODE_INCIDENT, "unknown mode %d", g->mode);
    CHECK(g->dft_accum == FDTD_DFT_COMPENSATED || g->dft_accum == FDTD_DFT_REFERENCE,
          "unknown dft_accum %d", g->dft_accum);
    {
        // the stencil kernels index every field, material and CPML array
        // with 32-bit offsets: each must hold fewer than 2^30 floats
        const long long pitch = pitch_of(g->nx), lim = 1LL << 30;
        const long long ncell = (long long)g->nz * g->ny * pitch;
        const long long psi_block = 4LL * g->npml * ((long long)g->ny * pitch + (long long)g->nz * g->ny +
                                                     (long long)g->nz * pitch);
        CHECK(ncell < lim, "grid %d x %d x %d holds %lld cells per field array; the limit is 2^30",
              g->nx, g->ny, g->nz, ncell);
        CHECK(psi_block < lim, "CPML state of %lld floats per field exceeds 2^30", psi_block);
    }
    if (g->mode == FDTD_MODE_MAIN) {
        CHECK(g->tfsf_z >= 0 && g->tfsf_z < g->nz, "tfsf_z %d outside [0, %d)", g->tfsf_z, g->nz);
    } else {
        CHECK(g->tfsf_z >= 0 && g->tfsf_z < g->nz, "feed-in plane %d outside [0, %d)", g->tfsf_z, g->nz);
        CHECK(g->read_out >= 1 && g->read_out < g->nz, "read-out plane %d outside [1, %d)", g->read_out, g->nz);
    }
    return FDTD_OK;
}

struct Sizes {
    long long pitch, ncell, psi_len[24], psi_total, nsamp, acc_total;
    int nfaces;
    int kind[kMaxFaces], pos[kMaxFaces], na[kMaxFaces], nb[kMaxFaces];
    long long face_off[kMaxFaces + 1];
};

static Sizes sizes_of(const FdtdGeom *g) {
    Sizes s{};
    const long long nx = g->nx, ny = g->ny, nz = g->nz, np = g->npml;
    s.pitch = pitch_of(g->nx);
    s.ncell = nz * ny * s.pitch;
    const long long du = np * ny * s.pitch, lr = nz * ny * np, fb = nz * np * s.pitch;
    for (int k = 0; k < 24; ++k) {
        const int face = (k % 12) / 2;   // d u l r f b
        s.psi_len[k] = (face < 2) ? du : (face < 4) ? lr : fb;
        s.psi_total += s.psi_len[k];
    }
    const int nxp = g->nx - 2 * g->npml, nyp = g->ny - 2 * g->npml, nzp = g->nz - 2 * g->npml;
    if (g->mode == FDTD_MODE_MAIN) {
        s.nfaces = 6;
        const int kinds[6] = {kFaceZ, kFaceZ, kFaceX, kFaceX, kFaceY, kFaceY};
        const int pos[6] = {g->npml + 1, g->nz - g->npml - 1, g->npml + 1, g->nx - g->npml - 1,
                            g->npml + 1, g->ny - g->npml - 1};
        const int na[6] = {nyp, nyp, nzp, nzp, nzp, nzp};
        const int nb[6] = {nxp, nxp, nyp, nyp, nxp, nxp};
        for (int f = 0; f < 6; ++f) {
            s.kind[f] = kinds[f]; s.pos[f] = pos[f]; s.na[f] = na[f]; s.nb[f] = nb[f];
        }
    } else {
        s.nfaces = 1;
        s.kind[0] = kFaceZ; s.pos[0] = g->read_out; s.na[0] = nyp; s.nb[0] = nxp;
    }
    s.face_off[0] = 0;
    for (int f = 0; f < s.nfaces; ++f) s.face_off[f + 1] = s.face_off[f] + (long long)s.na[f] * s.nb[f];
    s.nsamp = s.face_off[s.nfaces];
    s.acc_total = 4 * s.nsamp;
    (void)nx;
    return s;
}

// MAIN simulations hold two buffer sets of fields and H-CPML state (the
// fused stepper ping-pongs between them); INCIDENT ones one set.
static int nsets_of(const FdtdGeom *g) { return g->mode == FDTD_MODE_MAIN ? 2 : 1; }

static long long psi_h_len(const Sizes &s) {
    long long n = 0;
    for (int k = 12; k < 24; ++k) n += s.psi_len[k];
    return n;
}

// float4s of a simulation's x-face taps (Desc::tapx).
static long long tap_len(const FdtdGeom *g) {
    return g->mode == FDTD_MODE_MAIN ? 2LL * 2 * g->nz * g->ny : 0;
}

static long long footprint_bytes(const FdtdGeom *g) {
    const Sizes s = sizes_of(g);
    const int ns = nsets_of(g);
    long long b = 6 * field_stride(s.ncell) * 4 * ns + (s.psi_total + (ns - 1) * psi_h_len(s)) * 4 + s.acc_total * 16;
    if (g->mode == FDTD_MODE_MAIN) b += 2 * s.ncell * 4 + tap_len(g) * 16;
    b += (long long)(g->nt + 1) * (32 + 32) + 4096;
    b += (long long)std::min<size_t>(32u << 20, (size_t)g->nx * g->ny * g->nz * sizeof(float));  // upload staging
    return b;
}

// Twiddles of the running DFT for step t: exp(-1j*omega*tt*dt) with the
// phase formed as numba forms it, theta = (-omega * tt) * dt
// (cpu/field_updates.py:387,395).  MAIN: E at tt = t, H at tt = t - 0.5;
// INCIDENT swaps them (phasors_incident, :478,486).
static void make_twiddles(const FdtdGeom *g, const FdtdCoeffs *c, std::vector<double4> &w64,
                          std::vector<float4> &w32) {
    const int n = g->nt > 0 ? g->nt : 1;
    w64.resize(n);
    w32.resize(2 * (size_t)n);
    auto split = [](double v, float &hi, float &lo) {
        hi = (float)v;
        lo = (float)(v - (double)hi);
    };
    for (int t = 0; t < n; ++t) {
        double te = (double)t, th = (double)t - 0.5;
        if (g->mode == FDTD_MODE_INCIDENT) std::swap(te, th);
        const double pe = (-c->omega * te) * c->dt;
        const double ph = (-c->omega * th) * c->dt;
        const double ce = std::cos(pe), se = std::sin(pe), ch = std::cos(ph), sh = std::sin(ph);
        w64[t] = make_double4(ce, se, ch, sh);
        float4 a, b;
        split(0.5 * ce, a.x, a.y);
        split(0.5 * se, a.z, a.w);
        split(0.25 * ch, b.x, b.y);
        split(0.25 * sh, b.z, b.w);
        w32[2 * (size_t)t] = a;
        w32[2 * (size_t)t + 1] = b;
    }
}

// ---------------------------------------------------------------------------
// groups (launch sequencing)
// ---------------------------------------------------------------------------
static void group_clear_graphs(Group &G) {
    for (auto &kv : G.graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : G.graph_src) (void)hipGraphDestroy(kv.second);
    G.graphs.clear();
    G.graph_src.clear();
}

// Widest lane (4, 2 or 1 x cells) whose tiles leave at most 30 % of the
// lanes idle over the members' rows (x padding to whole tiles), among the
// widths every member's row pitch holds.
static bool tiles_fill(const std::vector<FdtdSim *> &members, int v, int tw) {
    bool fits = true;
    double used = 0, alloc = 0;
    for (const FdtdSim *s : members) {
        fits = fits && s->h.sy % v == 0;
        const double rows = (double)s->g.ny * s->g.nz;
        used += rows * s->g.nx;
        alloc += rows * (double)((s->g.nx + tw - 1) / tw * tw);
    }
    return fits && alloc <= 1.3 * used;
}

static int auto_xvec(const std::vector<FdtdSim *> &members, int xw) {
    for (int v : {4, 2})
        if (tiles_fill(members, v, kWave * v * xw)) return v;
    return 1;
}

// Whether flat tiles (kXwFlat) hold the members' cells in at most 95 % of
// the lane slots of row tiles xw_of(member) wavefronts wide (one cell per
// lane): narrow rows padded to whole wavefronts leave lanes idle that flat
// tiles, whose wavefronts run on into the next row, fill.  Every member's
// row pitch must hold a wavefront (sy >= 64).
template <typename XwOf>
static bool flat_pays(const std::vector<FdtdSim *> &members, XwOf xw_of) {
    double rows_slots = 0, flat_slots = 0;
    for (FdtdSim *s : members) {
        if (s->h.sy < kWave) return false;
        const long long tw = (long long)kWave * xw_of(s);
        rows_slots += (double)s->g.nz * s->g.ny * ((s->g.nx + tw - 1) / tw * tw);
        flat_slots += (double)s->g.nz * (((long long)s->g.ny * s->h.sy + 4 * kWave - 1) / (4 * kWave) * 4 * kWave);
    }
    return flat_slots <= 0.95 * rows_slots;
}

// Widest block (4 or 2 wavefronts along x, else 1) whose tiles of
// 64 xvec xw cells leave at most 30 % of the lanes idle.
static int auto_xwaves(const std::vector<FdtdSim *> &members, int v) {
    for (int w : {4, 2})
        if (tiles_fill(members, v, kWave * v * w)) return w;
    return 1;
}

// Slab size of the split stepper's interleaved H/E launches
// (FDTD_SLAB_BYTES; 0, the default, launches each kernel over the whole
// group).  Measured at 761 x 761 x 203: 256 MB slabs of 4-plane chunks gain
// about 4 %; other sizes lose more to the extra launches than the
// memory-side cache saves.
static long long slab_bytes() {
    const char *e = getenv("FDTD_SLAB_BYTES");   // read at each group build
    return e ? std::max(0LL, atoll(e)) : 0LL;
}

static int group_build(Group &G) {
    CHECK(!G.members.empty(), "empty group");
    HIP_TRY(hipSetDevice(G.device));
    std::vector<int4> tiles;
    std::vector<Desc> descs;
    G.mode = G.members[0]->g.mode;
    G.fused = G.members[0]->h.fused;
    const int xv = G.launch.xvec;
    int ty = G.launch.ty, xw = G.launch.xw;
    bool per_member = false;
    G.flat = false;
    CHECK(G.members.size() < (1u << 16), "a group holds at most 65535 simulations");
    CHECK(xv == 0 || xv == 1 || xv == 2 || xv == 4, "xvec must be 0, 1, 2 or 4 (got %d)", xv);
    CHECK(xw == 0 || xw == 1 || xw == 2 || xw == 4 || xw == kXwFlat,
          "xwaves must be 0, 1, 2, 4 or -1 (flat tiles) (got %d)", xw);
    if (xw == kXwFlat) {
        CHECK(G.members[0]->h.fused == 0, "flat tiles are a split-stepper shape");
        CHECK(xv == 0 || xv == 1, "flat tiles hold one x cell per lane (xvec 0 or 1, got %d)", xv);
        CHECK(ty == 0, "flat tiles take no ty (got %d)", ty);
        CHECK(G.launch.xcd == 0 || G.launch.xcd >= kXcdFine, "flat tiles run in plain order or fine bands (got %d)",
              G.launch.xcd);
        for (FdtdSim *s : G.members)
            CHECK(s->h.sy >= kWave, "flat tiles need a row pitch of at least %d cells (nx = %d)", kWave, s->g.nx);
    }
    CHECK(G.launch.xcd >= 0 && G.launch.xcd <= kXcdFine + 3,
          "block order must be 0 (plain), 1 (XCD-contiguous), 2 (banded) or 3-6 (fine bands of 1-8 row blocks) "
          "(got %d)", G.launch.xcd);
    const int band = G.launch.xcd >= kXcdFine ? 1 << (G.launch.xcd - kXcdFine) : 0;
    if (G.fused == 2) {
        CHECK(ty == 3 || ty == 7, "row-parallel stepper: ty (rows per block) must be 3 or 7 (got %d)", ty);
        CHECK(xw <= 1, "row-parallel stepper: xwaves must be 1");
        CHECK(xv == 0 || xv == 1 || xv == 2, "row-parallel stepper: xvec must be 1 or 2 (got %d)", xv);
        CHECK(G.launch.xcd < 2, "row-parallel stepper: block order must be 0 (plain) or 1 (XCD-contiguous)");
        xw = 1;
        G.xvec = xv ? xv : 1;
    } else if (G.fused == 1) {
        CHECK(ty == 2 || ty == 4 || ty == 8, "fused stepper: ty (rows per wavefront) must be 2, 4 or 8 (got %d)",
              ty);
        CHECK(xw <= 1, "fused stepper: one wavefront per strip (xwaves must be 1)");
        CHECK(G.launch.xcd < 2, "fused stepper: block order must be 0 (plain) or 1 (XCD-contiguous)");
        xw = 1;
        G.xvec = xv ? xv : auto_xvec(G.members, 1);
    } else {
        // the halo stepper's blocks may hold up to 8 wavefronts
        const int max_waves = G.fused == 3 ? 8 : 4;
        CHECK(ty == 0 || ty == 1 || ty == 2 || ty == 4 || (ty == 8 && max_waves == 8),
              "split stepper: ty must be 0, 1, 2 or 4 (halo stepper: also 8) (got %d)", ty);
        // automatic: one cell per lane unless the block width is fixed
        G.xvec = xv ? xv : xw ? auto_xvec(G.members, xw) : 1;
        if (G.fused == 3) {   // halo stepper: the split shapes, 1 or 2 cells per lane
            CHECK(xv == 0 || xv == 1 || xv == 2, "halo stepper: xvec must be 1 or 2 (got %d)", xv);
            G.xvec = std::min(G.xvec, 2);
        }
        // a batch with a fully automatic shape gives each member its own
        // block width (4-wavefront blocks, 4 / xw rows); its tiles carry it
        per_member = !xw && !ty && !xv && G.members.size() > 1;
        // a fully automatic split-stepper shape runs on flat tiles where
        // they fill the lanes better (flat_pays), for one simulation or a
        // batch (then every member flat)
        const bool automatic = !xw && !ty && !xv && G.launch.xcd == 0;
        if (xw == kXwFlat ||
            (automatic && G.fused == 0 &&
             flat_pays(G.members, [](FdtdSim *s) { return auto_xwaves(std::vector<FdtdSim *>{s}, 1); }))) {
            G.flat = true;
            per_member = false;
            G.xvec = 1;
            xw = 4;
            ty = 1;
        }
        if (per_member) {
            xw = 4;
            ty = 1;
        }
        if (!xw) xw = auto_xwaves(G.members, G.xvec);
        if (!ty) ty = std::max(1, 4 / xw);
        CHECK(xw * ty <= max_waves, "split stepper: xwaves x ty = %d exceeds %d wavefronts per block", xw * ty,
              max_waves);
    }
    G.ty = ty;
    G.xw = xw;
    if (G.fused == 1) {
        // a block's four wavefronts keep (4 ty + 2) x 64 xvec floats each in
        // LDS; the built shapes stay within 64 KB per block
        if (!xv) G.xvec = std::min(G.xvec, ty == 2 ? 4 : ty == 4 ? 2 : 1);
        CHECK((4 * ty + 2) * G.xvec <= 64, "fused stepper: %d rows of %d-cell lanes exceed the LDS budget "
              "(rows x cells: 2 x 4, 4 x 2, 8 x 1 at most)", ty, G.xvec);
    }
    for (FdtdSim *s : G.members)
        CHECK(s->h.sy % G.xvec == 0, "row pitch %lld does not hold whole %d-cell lanes", s->h.sy, G.xvec);
    // tiles of (64 xvec xw) x ty cells: split, xw wavefronts along each of
    // ty rows of a block; fused, one wavefront per tile, four tiles per block
    const int tx = kWave * G.xvec * xw;
    const int tys = ty;
    G.max_dft_items = 0;
    for (size_t m = 0; m < G.members.size(); ++m) {
        FdtdSim *s = G.members[m];
        // split tiles carry log2(xwaves) above the member index
        const int xw_m = per_member ? auto_xwaves(std::vector<FdtdSim *>{s}, G.xvec) : xw;
        const int lxw_m = G.fused == 1 || G.fused == 2 ? 0 : xw_m == 4 ? 2 : xw_m == 2 ? 1 : 0;
        const int tx_m = kWave * G.xvec * xw_m, ty_m = per_member ? 4 / xw_m : tys;
        const int tag = (int)m | (lxw_m << 16);
        CHECK(s->g.mode == G.mode, "a group cannot mix MAIN and INCIDENT simulations");
        CHECK(s->h.fused == G.fused, "a group cannot mix fused and split steppers");
        CHECK(s->device == G.device, "group members must share a device");
        descs.push_back(s->h);
        const int nz = s->g.nz, zc = G.launch.zchunk > 0 ? G.launch.zchunk : nz;
        std::vector<int4> chunk;
        if (G.flat) {
            // flat tiles: blocks of 4 wavefronts over 256 consecutive cells
            // of the plane's row-major (y, x) index
            const long long span = (long long)s->g.ny * s->h.sy;
            CHECK(span < (1LL << 31), "plane of %lld cells too large for flat tiles", span);
            const int nfb = (int)((span + 4 * kWave - 1) / (4 * kWave));
            for (int z0 = 0; z0 < nz; z0 += zc) {
                const int zw = z0 | (std::min(nz, z0 + zc) << 16);
                if (!band) {
                    for (int b = 0; b < nfb; ++b) tiles.push_back(make_int4(tag, b * 4 * kWave, 0, zw));
                    continue;
                }
                // fine bands over the flat blocks (the row blocks of one
                // column): slot i of a panel runs on XCD i % kXcds and takes
                // block (p kXcds + i % kXcds) band + i / kXcds; slots past
                // the last block are empty tiles
                const int per_panel = kXcds * band, panels = (nfb + per_panel - 1) / per_panel;
                for (int p = 0; p < panels; ++p)
                    for (int i = 0; i < per_panel; ++i) {
                        const int b = (p * kXcds + i % kXcds) * band + i / kXcds;
                        tiles.push_back(b < nfb ? make_int4(tag, b * 4 * kWave, 0, zw) : make_int4(0, 0, 0, 0));
                    }
            }
            G.max_dft_items = std::max(G.max_dft_items, s->h.dft_items);
            continue;
        }
        for (int z0 = 0; z0 < nz; z0 += zc) {
            const int z1 = std::min(nz, z0 + zc);
            chunk.clear();
            for (int y0 = 0; y0 < s->g.ny; y0 += ty_m)
                for (int x0 = 0; x0 < s->g.nx; x0 += tx_m)
                    chunk.push_back(make_int4(tag, x0, y0, z0 | (z1 << 16)));
            if (band) {
                // fine bands (the tile grid's order, TileGrid): panels of
                // kXcds bands of `band` row blocks; slot i of a panel runs on
                // XCD i % kXcds and takes the (i / kXcds)-th tile of its band,
                // x fastest.  Slots past the last row block are empty tiles.
                const int ntx = (s->g.nx + tx_m - 1) / tx_m, nyb = (s->g.ny + ty_m - 1) / ty_m;
                const int per_panel = kXcds * band * ntx, panels = (nyb + kXcds * band - 1) / (kXcds * band);
                for (int p = 0; p < panels; ++p)
                    for (int i = 0; i < per_panel; ++i) {
                        const int kk = i / kXcds, q = kk / ntx;
                        const int yb = (p * kXcds + i % kXcds) * band + q;
                        tiles.push_back(yb < nyb ? chunk[(size_t)yb * ntx + (kk - q * ntx)] : make_int4(0, 0, 0, 0));
                    }
                continue;
            }
            if (G.launch.xcd != 2) {
                tiles.insert(tiles.end(), chunk.begin(), chunk.end());
                continue;
            }
            // banded order: the dispatcher deals blocks to the XCDs in
            // turn, so block 8 k + j runs on XCD j; XCD j takes the k-th
            // tile of band j, the j-th eighth of the chunk's rows.  The
            // rows a tile's y +- 1 neighbours lie on then pass through one
            // XCD's L2, while all XCDs stream through the same z chunk.
            // Short bands are padded with empty tiles (w = 0).
            const int T = (int)chunk.size(), B = (T + kXcds - 1) / kXcds;
            for (int k = 0; k < B; ++k)
                for (int j = 0; j < kXcds; ++j) {
                    const int i = j * B + k;
                    tiles.push_back(i < T ? chunk[i] : make_int4(0, 0, 0, 0));
                }
        }
        G.max_dft_items = std::max(G.max_dft_items, s->h.dft_items);
    }
    group_clear_graphs(G);
    // a device free waits for all work on the device, so the tile list and
    // descriptors keep their allocations while they are large enough: a
    // rebuild while other streams run (a serving loop's admissions) then
    // does not wait for them
    if (descs.size() > G.descs_cap) {
        (void)hipFree(G.d_descs);
        G.d_descs = nullptr;
        G.descs_cap = 0;
        HIP_TRY(hipMalloc(&G.d_descs, descs.size() * sizeof(Desc)));
        G.descs_cap = descs.size();
    }
    if (tiles.size() > G.tiles_cap) {
        (void)hipFree(G.d_tiles);
        G.d_tiles = nullptr;
        G.tiles_cap = 0;
        HIP_TRY(hipMalloc(&G.d_tiles, tiles.size() * sizeof(int4)));
        G.tiles_cap = tiles.size();
    }
    HIP_TRY(hipMemcpyAsync(G.d_descs, descs.data(), descs.size() * sizeof(Desc), hipMemcpyHostToDevice,
                           G.stream));
    HIP_TRY(hipMemcpyAsync(G.d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice,
                           G.stream));
    HIP_TRY(hipStreamSynchronize(G.stream));
    G.ntiles = (int)tiles.size();
    G.nmodels = (int)descs.size();
    G.tg = TileGrid{};
    G.step1 = nullptr;
    if (G.members.size() == 1 && G.flat) {
        const FdtdSim *m = G.members[0];
        const int zc = G.launch.zchunk > 0 ? G.launch.zchunk : m->g.nz;
        const long long span = (long long)m->g.ny * m->h.sy;
        const int nfb = (int)((span + 4 * kWave - 1) / (4 * kWave));
        G.tg = band ? TileGrid{1, nfb, 4 * kWave, 0, zc, m->g.nz, 0, band, 1}
                    : TileGrid{nfb, 1, 4 * kWave, 0, zc, m->g.nz, 0, 0, 1};
        CHECK((long long)tilegrid_gx(G.tg) * tilegrid_gy(G.tg) * ((m->g.nz + zc - 1) / zc) == G.ntiles,
              "tile grid mismatch");
        G.step1 = m->h.step;
    } else if (G.members.size() == 1 && (G.fused == 0 || G.fused == 3) && (G.launch.xcd == 0 || band)) {
        const FdtdSim *m = G.members[0];
        const int zc = G.launch.zchunk > 0 ? G.launch.zchunk : m->g.nz;
        G.tg = TileGrid{(m->g.nx + tx - 1) / tx, (m->g.ny + tys - 1) / tys, tx, tys, zc, m->g.nz, 0, band, 0};
        CHECK((long long)tilegrid_gx(G.tg) * tilegrid_gy(G.tg) * ((m->g.nz + zc - 1) / zc) == G.ntiles,
              "tile grid mismatch");
        G.step1 = m->h.step;
    }
    if (G.tg.ntx > 0 && tilegrid_gy(G.tg) > 65535) {
        // more row blocks than a launch grid's y extent holds (very many
        // narrow rows): the kernels take their tiles from the list instead
        // (a 1-D grid), built above in the same order
        G.tg = TileGrid{};
        G.step1 = nullptr;
    }
    G.slabs.assign(1, {0, G.ntiles});
    // A single MAIN simulation on the split stepper runs the H and E kernels
    // slab by slab (whole z chunks, about slab_bytes of field and material
    // state each): E of a slab then reads H, and E itself, while they are
    // still in the 256 MB memory-side cache, which takes the E kernel's
    // re-reads off HBM.  H of slab s needs only E at z and z+1 (old), E of
    // slab s needs H at z-1 (new, from slab s-1), so this order is exact.
    const long long sb = slab_bytes();
    if (sb > 0 && !G.fused && G.mode == FDTD_MODE_MAIN && G.members.size() == 1) {
        const FdtdSim *m = G.members[0];
        const int zc = G.launch.zchunk > 0 ? G.launch.zchunk : m->g.nz;
        int per_chunk = (int)(((m->g.ny + tys - 1) / tys) * ((m->g.nx + tx - 1) / tx));
        if (G.launch.xcd == 2) per_chunk = (per_chunk + kXcds - 1) / kXcds * kXcds;
        if (band) {
            const TileGrid tg{(m->g.nx + tx - 1) / tx, (m->g.ny + tys - 1) / tys, tx, tys, zc, m->g.nz, 0, band, 0};
            per_chunk = tilegrid_gx(tg) * tilegrid_gy(tg);
        }
        if (G.flat) {
            const int nfb = (int)(((long long)m->g.ny * m->h.sy + 4 * kWave - 1) / (4 * kWave));
            per_chunk = band ? (nfb + kXcds * band - 1) / (kXcds * band) * (kXcds * band) : nfb;
        }
        const long long chunk_bytes = 32LL * zc * m->g.ny * m->g.nx;
        const int chunks = (int)std::max(1LL, sb / chunk_bytes);
        G.slabs.clear();
        for (int t0 = 0; t0 < G.ntiles; t0 += chunks * per_chunk)
            G.slabs.push_back({t0, std::min(G.ntiles, t0 + chunks * per_chunk)});
    }
    return FDTD_OK;
}

static void group_free(Group &G) {
    (void)hipSetDevice(G.device);
    group_clear_graphs(G);
    (void)hipFree(G.d_descs);
    (void)hipFree(G.d_tiles);
    if (G.stream) (void)hipStreamDestroy(G.stream);
    G = Group();
}

enum KernelId { KH = 0, KE = 1, KD = 2, KF = 3 };

template <typename K>
static void launch_split(K direct, K listed, bool is_direct, dim3 grid, dim3 blk, Group &G, hipStream_t st,
                         const int4 *tl, int nt, int shape, const TileGrid &tg) {
    hipLaunchKernelGGL(is_direct ? direct : listed, grid, blk, 0, st, G.d_descs, tl, nt, shape, tg, G.step1);
}

// One kernel of a step of group G on stream st (G's own stream unless a
// batch runs the group's launches on its stream).
static void launch_kernel(Group &G, KernelId k, int t0 = 0, int t1 = -1, hipStream_t st = nullptr) {
    if (!st) st = G.stream;
    const dim3 blk(kWave, G.ty * G.xw);
    const int lxw = G.xw == 4 ? 2 : G.xw == 2 ? 1 : 0;
    const int shape = lxw | (G.launch.xcd == 1 ? kShapeXcd : 0);
    if (t1 < 0) t1 = G.ntiles;
    const int4 *tl = G.d_tiles + t0;
    const int nt = t1 - t0;
    // split kernels of a single simulation: a 3-D grid over tiles [t0, t1),
    // which are whole z chunks
    TileGrid tg = G.tg;
    dim3 sgrid((unsigned)nt);
    if (tg.ntx > 0) {
        const int gx = tilegrid_gx(tg), gy = tilegrid_gy(tg), per = gx * gy;
        tg.t0 = t0 / per;
        sgrid = dim3((unsigned)gx, (unsigned)gy, (unsigned)(nt / per));
    }
    if (k == KF && G.fused == 3) {
        if (G.xvec == 2)
            launch_split(k_step_halo<2, true>, k_step_halo<2, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt, shape,
                         tg);
        else
            launch_split(k_step_halo<1, true>, k_step_halo<1, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt, shape,
                         tg);
    } else if (k == KF && G.fused == 2) {
        const int r = G.ty, v = G.xvec;
        auto *fn = v == 2 ? (r == 7 ? &k_step_rows<2, 7> : &k_step_rows<2, 3>)
                          : (r == 7 ? &k_step_rows<1, 7> : &k_step_rows<1, 3>);
        hipLaunchKernelGGL(fn, dim3(G.ntiles), dim3(kWave, r + 1), 0, st, G.d_descs, G.d_tiles, G.ntiles,
                           shape);
    } else if (k == KF) {
        const dim3 grid((G.ntiles + 3) / 4), b4(kWave, 4);
        auto *fn = &k_step_strip<1, 2>;
        const int r = G.ty, v = G.xvec;
        if (v == 4) fn = &k_step_strip<4, 2>;
        else if (v == 2) fn = r == 4 ? &k_step_strip<2, 4> : &k_step_strip<2, 2>;
        else fn = r == 8 ? &k_step_strip<1, 8> : r == 4 ? &k_step_strip<1, 4> : &k_step_strip<1, 2>;
        hipLaunchKernelGGL(fn, grid, b4, 0, st, G.d_descs, G.d_tiles, G.ntiles, shape);
    } else if (k == KH) {
        if (G.flat)
            launch_split(k_h_half<1, true, true>, k_h_half<1, false, true>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else if (G.xvec == 4)
            launch_split(k_h_half<4, true, false>, k_h_half<4, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else if (G.xvec == 2)
            launch_split(k_h_half<2, true, false>, k_h_half<2, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else
            launch_split(k_h_half<1, true, false>, k_h_half<1, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
    } else if (k == KE) {
        if (G.flat)
            launch_split(k_e_half<1, true, true>, k_e_half<1, false, true>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else if (G.xvec == 4)
            launch_split(k_e_half<4, true, false>, k_e_half<4, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else if (G.xvec == 2)
            launch_split(k_e_half<2, true, false>, k_e_half<2, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
        else
            launch_split(k_e_half<1, true, false>, k_e_half<1, false, false>, tg.ntx > 0, sgrid, blk, G, st, tl, nt,
                         shape, tg);
    } else {
        const unsigned gx = (unsigned)std::max(1LL, (G.max_dft_items + 255) / 256);
        hipLaunchKernelGGL(k_dft, dim3(gx, G.nmodels), dim3(256), 0, st, G.d_descs);
    }
}

static void launch_step(Group &G, hipStream_t st = nullptr) {
    if (G.mode == FDTD_MODE_MAIN && G.fused) {   // H then E in one kernel, phasors
        launch_kernel(G, KF, 0, -1, st);
    } else if (G.mode == FDTD_MODE_MAIN) {       // cpu/main_sim.py:140-195: H, E, phasors
        for (const auto &sl : G.slabs) {
            launch_kernel(G, KH, sl.first, sl.second, st);
            launch_kernel(G, KE, sl.first, sl.second, st);
        }
    } else {                          // cpu/incident_sim.py:134-161: E, H, record + phasors
        launch_kernel(G, KE, 0, -1, st);
        launch_kernel(G, KH, 0, -1, st);
    }
    launch_kernel(G, KD, 0, -1, st);
}

static int group_run(Group &G, int nsteps, bool use_graph) {
    CHECK(nsteps >= 0, "negative step count");
    if (nsteps == 0) return FDTD_OK;
    HIP_TRY(hipSetDevice(G.device));
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) launch_step(G);
        HIP_TRY(hipGetLastError());
        return FDTD_OK;
    }
    // Replay a captured chunk of steps; the chunk length is bounded so graph
    // instantiation stays cheap, and the remainder runs as its own graph.
    const int chunk = std::min(nsteps, 64);
    auto replay = [&](int len) -> int {
        auto it = G.graphs.find(len);
        if (it == G.graphs.end()) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < len; ++s) launch_step(G);
            HIP_TRY(hipStreamEndCapture(G.stream, &graph));
            hipGraphExec_t exec;
            HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            G.graph_src[len] = graph;
            it = G.graphs.emplace(len, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, G.stream));
        return FDTD_OK;
    };
    int done = 0;
    while (done + chunk <= nsteps) {
        int rc = replay(chunk);
        if (rc) return rc;
        done += chunk;
    }
    if (done < nsteps) {
        int rc = replay(nsteps - done);
        if (rc) return rc;
    }
    return FDTD_OK;
}

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
extern "C" int fdtd_abi_version(void) { return FDTD_ABI_VERSION; }
extern "C" const char *fdtd_last_error(void) { return g_err.c_str(); }

static int n2f_scratch_ptr(int device, void **out);

extern "C" int fdtd_init(int device) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    CHECK(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    void *scratch_ptr = nullptr;   // the no-allocation entry point's work area on this device
    return n2f_scratch_ptr(device, &scratch_ptr);
}

extern "C" int fdtd_device_info(FdtdDeviceInfo *out) {
    CHECK(out, "out is NULL");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, g_device));
    HIP_TRY(hipSetDevice(g_device));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    memset(out, 0, sizeof *out);
    out->device = g_device;
    out->cu_count = p.multiProcessorCount;
    out->xcd_count = kXcds;
    out->warp_size = p.warpSize;
    out->total_mem = (int64_t)tot;
    out->free_mem = (int64_t)fr;
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    snprintf(out->name, sizeof out->name, "%s", p.name);
    return FDTD_OK;
}

// Plain streaming copy between two fresh device buffers of nbytes each,
// timed with HIP events on a stream of its own (default priority).
extern "C" int fdtd_probe_copy(int64_t nbytes, int vec, int reps, float *ms_per_copy) {
    CHECK(nbytes >= 1024 && nbytes % 16 == 0, "nbytes must be a multiple of 16, at least 1024");
    CHECK(vec == 1 || vec == 4, "vec must be 1 or 4");
    CHECK(reps >= 1 && ms_per_copy, "bad reps / output");
    HIP_TRY(hipSetDevice(g_device));
    const long long n = nbytes / 4;
    const long long threads = vec == 4 ? n / 4 : n;
    const long long blocks = (threads + 255) / 256;
    CHECK(blocks <= 0x7fffffffLL, "copy too large for one launch");
    float *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = FDTD_OK;
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == FDTD_OK) rc = fail(FDTD_EHIP, "%s: %s", what, hipGetErrorString(e));
        return rc == FDTD_OK;
    };
    if (hip_ok(hipMalloc(&src, nbytes), "hipMalloc") && hip_ok(hipMalloc(&dst, nbytes), "hipMalloc") &&
        hip_ok(hipStreamCreate(&st), "hipStreamCreate") && hip_ok(hipEventCreate(&ev[0]), "hipEventCreate") &&
        hip_ok(hipEventCreate(&ev[1]), "hipEventCreate") &&
        hip_ok(hipMemsetAsync(src, 0x3f, nbytes, st), "hipMemsetAsync")) {
        for (int r = 0; r <= reps && rc == FDTD_OK; ++r) {   // r = 0: warm-up
            if (r == 1) hip_ok(hipEventRecord(ev[0], st), "hipEventRecord");
            if (vec == 4)
                hipLaunchKernelGGL(k_copy_probe<4>, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
            else
                hipLaunchKernelGGL(k_copy_probe<1>, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
            hip_ok(hipGetLastError(), "k_copy_probe");
        }
        float ms = 0;
        if (hip_ok(hipEventRecord(ev[1], st), "hipEventRecord") &&
            hip_ok(hipEventSynchronize(ev[1]), "hipEventSynchronize") &&
            hip_ok(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime"))
            *ms_per_copy = ms / reps;
    }
    if (st) hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}

// nread input and nwrite output arrays of n floats each, one element of
// each per thread; timed like fdtd_probe_copy.
extern "C" int fdtd_probe_streams(int64_t n, int nread, int nwrite, int reps, float *ms_per_pass) {
    const bool known = (nread == 1 && nwrite == 1) || (nread == 6 && nwrite == 3) ||
                       (nread == 8 && nwrite == 3) || (nread == 8 && nwrite == 6);
    CHECK(known, "stream mix %d reads + %d writes not built (1+1, 6+3, 8+3, 8+6)", nread, nwrite);
    CHECK(n >= 256 && n <= (1LL << 31), "n outside [256, 2^31]");
    CHECK(reps >= 1 && ms_per_pass, "bad reps / output");
    HIP_TRY(hipSetDevice(g_device));
    float *in = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = FDTD_OK;
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == FDTD_OK) rc = fail(FDTD_EHIP, "%s: %s", what, hipGetErrorString(e));
        return rc == FDTD_OK;
    };
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (hip_ok(hipMalloc(&in, (size_t)nread * n * sizeof(float)), "hipMalloc") &&
        hip_ok(hipMalloc(&out, (size_t)nwrite * n * sizeof(float)), "hipMalloc") &&
        hip_ok(hipStreamCreate(&st), "hipStreamCreate") && hip_ok(hipEventCreate(&ev[0]), "hipEventCreate") &&
        hip_ok(hipEventCreate(&ev[1]), "hipEventCreate") &&
        hip_ok(hipMemsetAsync(in, 0, (size_t)nread * n * sizeof(float), st), "hipMemsetAsync")) {
        for (int r = 0; r <= reps && rc == FDTD_OK; ++r) {   // r = 0: warm-up
            if (r == 1) hip_ok(hipEventRecord(ev[0], st), "hipEventRecord");
            if (nread == 1)
                hipLaunchKernelGGL((k_stream_probe<1, 1>), dim3(blocks), dim3(256), 0, st, in, out, (long long)n);
            else if (nread == 6)
                hipLaunchKernelGGL((k_stream_probe<6, 3>), dim3(blocks), dim3(256), 0, st, in, out, (long long)n);
            else if (nwrite == 3)
                hipLaunchKernelGGL((k_stream_probe<8, 3>), dim3(blocks), dim3(256), 0, st, in, out, (long long)n);
            else
                hipLaunchKernelGGL((k_stream_probe<8, 6>), dim3(blocks), dim3(256), 0, st, in, out, (long long)n);
            hip_ok(hipGetLastError(), "k_stream_probe");
        }
        float ms = 0;
        if (hip_ok(hipEventRecord(ev[1], st), "hipEventRecord") &&
            hip_ok(hipEventSynchronize(ev[1]), "hipEventSynchronize") &&
            hip_ok(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime"))
            *ms_per_pass = ms / reps;
    }
    if (st) hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    return rc;
}

extern "C" int fdtd_probe_stencil(int nx, int ny, int nz, int zchunk, int kind, int reps, float *ms_per_pass) {
    CHECK(kind >= 0 && kind <= 2, "kind must be 0 (H half), 1 (E half) or 2 (halo stepper)");
    CHECK(nx >= 1 && ny >= 1 && nz >= 1 && zchunk >= 1, "bad sizes");
    CHECK(reps >= 1 && ms_per_pass, "bad reps / output");
    const long long sy = pitch_of(nx), sz = sy * ny, ncell = sz * nz;
    CHECK(ncell < (1LL << 30), "grid of %lld cells exceeds 2^30", ncell);
    const long long stride = field_stride(ncell);
    const int ntx = (nx + 255) / 256;
    const long long blocks = (long long)ntx * ny * ((nz + zchunk - 1) / zchunk);
    CHECK(blocks < (1LL << 31), "too many blocks");
    HIP_TRY(hipSetDevice(g_device));
    float *base = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = FDTD_OK;
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == FDTD_OK) rc = fail(FDTD_EHIP, "%s: %s", what, hipGetErrorString(e));
        return rc == FDTD_OK;
    };
    const size_t bytes = (size_t)(kind == 2 ? 14 : 8) * stride * sizeof(float);
    if (hip_ok(hipMalloc(&base, bytes), "hipMalloc") && hip_ok(hipStreamCreate(&st), "hipStreamCreate") &&
        hip_ok(hipEventCreate(&ev[0]), "hipEventCreate") && hip_ok(hipEventCreate(&ev[1]), "hipEventCreate") &&
        hip_ok(hipMemsetAsync(base, 0, bytes, st), "hipMemsetAsync")) {
        for (int r = 0; r <= reps && rc == FDTD_OK; ++r) {   // r = 0: warm-up
            if (r == 1) hip_ok(hipEventRecord(ev[0], st), "hipEventRecord");
            if (kind == 0)
                hipLaunchKernelGGL(k_stencil_probe<0>, dim3((unsigned)blocks), dim3(256), 0, st, base, stride, nx,
                                   ny, nz, (uint32_t)sy, (uint32_t)sz, zchunk, ntx);
            else if (kind == 2)
                hipLaunchKernelGGL(k_stencil_probe<2>, dim3((unsigned)blocks), dim3(256), 0, st, base, stride, nx,
                                   ny, nz, (uint32_t)sy, (uint32_t)sz, zchunk, ntx);
            else
                hipLaunchKernelGGL(k_stencil_probe<1>, dim3((unsigned)blocks), dim3(256), 0, st, base, stride, nx,
                                   ny, nz, (uint32_t)sy, (uint32_t)sz, zchunk, ntx);
            hip_ok(hipGetLastError(), "k_stencil_probe");
        }
        float ms = 0;
        if (hip_ok(hipEventRecord(ev[1], st), "hipEventRecord") &&
            hip_ok(hipEventSynchronize(ev[1]), "hipEventSynchronize") &&
            hip_ok(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime"))
            *ms_per_pass = ms / reps;
    }
    if (st) hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
    if (base) (void)hipFree(base);
    return rc;
}

// ---------------------------------------------------------------------------
// simulations
// ---------------------------------------------------------------------------
extern "C" int fdtd_sim_footprint(const FdtdGeom *g, int64_t *bytes) {
    int rc = check_geom(g);
    if (rc) return rc;
    CHECK(bytes, "bytes is NULL");
    *bytes = footprint_bytes(g);
    return FDTD_OK;
}

static int sim_alloc(FdtdSim *s) {
    const FdtdGeom *g = &s->g;
    const Sizes z = sizes_of(g);
    s->pitch = z.pitch;
    s->ncell = z.ncell;
    const int ns = nsets_of(g);
    const long long psi_alloc = z.psi_total + (ns - 1) * psi_h_len(z);
    HIP_TRY(hipMalloc(&s->fields, 6 * field_stride(z.ncell) * ns * sizeof(float)));
    if (g->mode == FDTD_MODE_MAIN) HIP_TRY(hipMalloc(&s->mat, 2 * z.ncell * sizeof(float)));
    HIP_TRY(hipMalloc(&s->psi, psi_alloc * sizeof(float)));
    HIP_TRY(hipMalloc(&s->acc, z.acc_total * sizeof(float4)));
    // initialisation is ordered on the simulation's own stream, so creating
    // a simulation does not wait for work other streams have queued
    hipStream_t st = s->self.stream;
    if (tap_len(g)) {
        HIP_TRY(hipMalloc(&s->tapx, tap_len(g) * sizeof(float4)));
        HIP_TRY(hipMemsetAsync(s->tapx, 0, tap_len(g) * sizeof(float4), st));
    }
    std::vector<double4> w64;
    std::vector<float4> w32;
    make_twiddles(g, &s->c, w64, w32);
    HIP_TRY(hipMalloc(&s->tw64, w64.size() * sizeof(double4)));
    HIP_TRY(hipMalloc(&s->tw32, w32.size() * sizeof(float4)));
    HIP_TRY(hipMemcpyAsync(s->tw64, w64.data(), w64.size() * sizeof(double4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->tw32, w32.data(), w32.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMalloc(&s->step, 2 * sizeof(int)));   // counter, step of the DFT
    // the field arrays' padding columns are read as x+1 neighbours never,
    // but keep all device memory defined
    HIP_TRY(hipMemsetAsync(s->fields, 0, 6 * field_stride(z.ncell) * ns * sizeof(float), st));
    if (s->mat) HIP_TRY(hipMemsetAsync(s->mat, 0, 2 * z.ncell * sizeof(float), st));
    HIP_TRY(hipStreamSynchronize(st));                // w64, w32 are host temporaries

    Desc &d = s->h;
    memset(&d, 0, sizeof d);
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz; d.npml = g->npml;
    d.tfsf_z = g->tfsf_z; d.read_out = g->read_out; d.nt = g->nt; d.nt_inc = g->nt_inc;
    d.mode = g->mode; d.dft_accum = g->dft_accum; d.uni_z = g->nz; d.n_src = 0;
    d.sy = z.pitch; d.sz = (long long)g->ny * z.pitch;
    d.du_sz = (long long)g->ny * z.pitch;
    d.lr_sz = (long long)g->ny * g->npml;
    d.fb_sz = (long long)g->npml * z.pitch;
    d.Db = s->c.Db; d.dtm0 = s->c.dtm0; d.dx = s->c.dx; d.dte0dx = s->c.dte0dx;
    memcpy(d.prof, s->c.prof, sizeof d.prof);
    for (int k = 0; k < 6; ++k) {
        d.f[0][k] = s->fields + k * field_stride(z.ncell);
        d.f[1][k] = s->fields + ((ns - 1) * 6 + k) * field_stride(z.ncell);
    }
    d.fused = 0;
    if (s->mat) {
        d.ca = s->mat;
        d.cb = s->mat + z.ncell;
    }
    long long off = 0;
    for (int k = 0; k < 24; ++k) {
        d.psi[0][k] = s->psi + off;
        s->psi_len[k] = z.psi_len[k];
        off += z.psi_len[k];
    }
    for (int k = 0; k < 24; ++k) {
        if (k < 12 || ns == 1) {
            d.psi[1][k] = d.psi[0][k];
        } else {
            d.psi[1][k] = s->psi + off;
            off += z.psi_len[k];
        }
    }
    d.acc = s->acc;
    for (int k = 0; k < 4 * z.nfaces; ++k) {
        const int f = k / 4;
        d.acc_off[k] = 4 * z.face_off[f] + (long long)(k % 4) * (z.face_off[f + 1] - z.face_off[f]);
    }
    // acc layout: face f occupies [4*face_off[f], 4*face_off[f+1]); within it,
    // component c starts at c * face_size.
    d.acc_off[24] = z.acc_total;
    d.nfaces = z.nfaces;
    for (int f = 0; f < z.nfaces; ++f) {
        d.face_kind[f] = z.kind[f]; d.face_pos[f] = z.pos[f];
        d.face_na[f] = z.na[f]; d.face_nb[f] = z.nb[f];
    }
    for (int f = 0; f <= z.nfaces; ++f) d.face_off[f] = z.face_off[f];
    d.dft_items = z.nsamp;
    d.tw64 = s->tw64;
    d.tw32 = s->tw32;
    d.step = s->step;
    d.tapx = s->tapx;
    return FDTD_OK;
}

static void sim_free(FdtdSim *s) {
    (void)hipSetDevice(s->device);
    group_free(s->self);
    (void)hipFree(s->fields);
    (void)hipFree(s->mat);
    (void)hipFree(s->psi);
    (void)hipFree(s->acc);
    (void)hipFree(s->tapx);
    (void)hipFree(s->tw64);
    (void)hipFree(s->tw32);
    (void)hipFree(s->src);
    (void)hipFree(s->step);
    (void)hipFree(s->scratch_c64);
    (void)hipFree(s->mat_stage);
    (void)hipFree(s->vol_stage);
    beam_unref(s->beam);
}

static int sim_refresh(FdtdSim *s) {
    // descriptor changed on the host: rebuild the batch-of-one
    return group_build(s->self);
}

extern "C" int fdtd_sim_create(const FdtdGeom *g, const FdtdCoeffs *c, FdtdSim **out) {
    int rc = check_geom(g);
    if (rc) return rc;
    CHECK(c && out, "NULL argument");
    *out = nullptr;
    FdtdSim *s = new FdtdSim();
    s->device = g_device;
    s->g = *g;
    s->c = *c;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) {
        delete s;
        return fail(FDTD_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
    }
    s->self.device = s->device;
    e = hipStreamCreateWithFlags(&s->self.stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(FDTD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    if (rc == FDTD_OK) rc = sim_alloc(s);
    if (rc == FDTD_OK) {
        s->self.members = {s};
        rc = group_build(s->self);
    }
    if (rc == FDTD_OK) rc = fdtd_sim_reset(s);
    if (rc != FDTD_OK) {
        std::string keep = g_err;
        sim_free(s);
        delete s;
        g_err = keep;
        return rc;
    }
    *out = s;
    return FDTD_OK;
}

extern "C" void fdtd_sim_destroy(FdtdSim *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->self.stream) (void)hipStreamSynchronize(s->self.stream);
    sim_free(s);
    delete s;
}

// Host (nz, ny, nx) <-> device (nz, ny, pitch).  Pageable host memory goes
// through a contiguous device staging buffer in bulk copies of whole rows;
// the pitch change is a device-to-device 2-D copy.  (A pitched copy straight
// from pageable memory is staged row by row by the runtime: ~1e5 small
// copies for a large grid.)  The staging buffer (at most kStageBytes) is the
// simulation's own and kept until it is destroyed: a device free waits for
// all work on the device, which would stall an upload made while other
// simulations run.
static constexpr size_t kStageBytes = 32u << 20;

static size_t stage_bytes(const FdtdGeom *g) {
    return std::min(kStageBytes, (size_t)g->nx * g->ny * g->nz * sizeof(float));
}

static int copy_volume_in(FdtdSim *s, float *dst, const float *src) {
    const size_t nx = (size_t)s->g.nx, rows = (size_t)s->g.ny * s->g.nz;
    if ((long long)nx == s->pitch) {
        HIP_TRY(hipMemcpyAsync(dst, src, nx * rows * sizeof(float), hipMemcpyHostToDevice, s->self.stream));
        return FDTD_OK;
    }
    const size_t want = stage_bytes(&s->g);
    if (s->vol_stage_len < want) {
        (void)hipFree(s->vol_stage);
        s->vol_stage = nullptr;
        s->vol_stage_len = 0;
        HIP_TRY(hipMalloc(&s->vol_stage, want));
        s->vol_stage_len = want;
    }
    const size_t per = std::max<size_t>(1, s->vol_stage_len / (nx * sizeof(float)));
    for (size_t r0 = 0; r0 < rows; r0 += per) {
        const size_t n = std::min(per, rows - r0);
        HIP_TRY(hipMemcpyAsync(s->vol_stage, src + r0 * nx, n * nx * sizeof(float), hipMemcpyHostToDevice,
                               s->self.stream));
        HIP_TRY(hipMemcpy2DAsync(dst + r0 * s->pitch, s->pitch * sizeof(float), s->vol_stage,
                                 nx * sizeof(float), nx * sizeof(float), n, hipMemcpyDeviceToDevice,
                                 s->self.stream));
        // the staging buffer is reused by the next rows
        HIP_TRY(hipStreamSynchronize(s->self.stream));
    }
    return FDTD_OK;
}

static int copy_rows_out(FdtdSim *s, float *dst, const float *src, size_t rows) {
    const size_t nx = (size_t)s->g.nx;
    if ((long long)nx == s->pitch) {
        HIP_TRY(hipMemcpyAsync(dst, src, nx * rows * sizeof(float), hipMemcpyDeviceToHost, s->self.stream));
        HIP_TRY(hipStreamSynchronize(s->self.stream));
        return FDTD_OK;
    }
    float *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, nx * rows * sizeof(float)));
    hipError_t e = hipMemcpy2DAsync(tmp, nx * sizeof(float), src, s->pitch * sizeof(float), nx * sizeof(float),
                                    rows, hipMemcpyDeviceToDevice, s->self.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst, tmp, nx * rows * sizeof(float), hipMemcpyDeviceToHost, s->self.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->self.stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(FDTD_EHIP, "volume download failed: %s", hipGetErrorString(e));
    return FDTD_OK;
}

static int copy_volume_out(FdtdSim *s, float *dst, const float *src) {
    return copy_rows_out(s, dst, src, (size_t)s->g.ny * s->g.nz);
}

extern "C" int fdtd_sim_set_materials(FdtdSim *s, const float *Ca, const float *Cb) {
    CHECK(s && Ca && Cb, "NULL argument");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "per-cell materials are for MAIN simulations");
    HIP_TRY(hipSetDevice(s->device));
    int rc = copy_volume_in(s, s->mat, Ca);
    if (!rc) rc = copy_volume_in(s, s->mat + s->ncell, Cb);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    // Top planes where every cell holds the same (Ca, Cb) -- the vacuum that
    // compute_material_arrays appends -- are read as scalars by k_e_half.
    const size_t plane = (size_t)s->g.ny * s->g.nx;
    const float ca0 = Ca[(size_t)(s->g.nz - 1) * plane], cb0 = Cb[(size_t)(s->g.nz - 1) * plane];
    int zu = s->g.nz;
    for (int z = s->g.nz - 1; z >= 0; --z) {
        const float *a = Ca + (size_t)z * plane, *b = Cb + (size_t)z * plane;
        bool same = true;
        for (size_t i = 0; i < plane && same; ++i) same = (a[i] == ca0) & (b[i] == cb0);
        if (!same) break;
        zu = z;
    }
    s->mat_set = true;
    s->h.uni_z = zu;
    s->h.ca_u = ca0;
    s->h.cb_u = cb0;
    return sim_refresh(s);
}

extern "C" int fdtd_sim_build_materials(FdtdSim *s, const FdtdMatSpec *spec, const int32_t *surfaces,
                                        const FdtdMatCylinder *cyl) {
    CHECK(s && spec, "NULL argument");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "per-cell materials are for MAIN simulations");
    const FdtdMatSpec &m = *spec;
    const int nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
    CHECK(m.nz_mat >= 0 && m.nz_mat <= nz, "nz_mat %d outside [0, %d]", m.nz_mat, nz);
    CHECK(m.nlayers >= 0 && m.nlayers <= FDTD_MAX_LAYERS, "nlayers %d outside [0, %d]", m.nlayers,
          FDTD_MAX_LAYERS);
    CHECK(m.nlayers == 0 || surfaces, "surfaces is NULL");
    CHECK(m.ncyl >= 0 && (m.ncyl == 0 || cyl), "bad cylinder list");
    CHECK(m.vox_z1 >= 0 && m.vox_z1 <= m.nz_mat, "vox_z1 %d outside [0, nz_mat]", m.vox_z1);
    CHECK(m.vox_frac >= 0.0 && m.vox_frac <= 1.0, "vox_frac outside [0, 1]");
    CHECK(m.dt > 0.0f && m.dx > 0.0f && m.eps0 > 0.0f, "dt, dx and eps0 must be positive");
    for (int k = 0; k < m.nlayers; ++k) {
        const FdtdMatLayer &L = m.layer[k];
        CHECK(L.kind == FDTD_LAYER_CONST || L.kind == FDTD_LAYER_RANDOM || L.kind == FDTD_LAYER_GRADIENT,
              "layer %d: unknown kind %d", k, L.kind);
        CHECK(L.kind != FDTD_LAYER_GRADIENT || k >= 1, "layer %d: a gradient layer needs a layer below", k);
        CHECK(L.kind != FDTD_LAYER_GRADIENT || (L.f0 < L.f1 && L.freq >= L.f0 && L.freq <= L.f1),
              "layer %d: frequency outside its table interval", k);
    }
    for (int c = 0; c < m.ncyl; ++c)
        CHECK(cyl[c].lo[0] >= 0 && cyl[c].lo[1] >= 0 && cyl[c].lo[2] >= 0 && cyl[c].hi[0] <= m.nz_mat &&
                  cyl[c].hi[1] <= ny && cyl[c].hi[2] <= nx && cyl[c].lo[0] <= cyl[c].hi[0] &&
                  cyl[c].lo[1] <= cyl[c].hi[1] && cyl[c].lo[2] <= cyl[c].hi[2],
              "cylinder %d: cell box outside the material volume", c);
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    const size_t plane = (size_t)nx * ny;
    // surface maps and cylinder list staged in one device buffer of the
    // simulation, grown when too small and kept until it is destroyed
    const size_t surf_bytes = (m.nlayers * plane * sizeof(int32_t) + 255) & ~(size_t)255;
    const size_t need = surf_bytes + m.ncyl * sizeof(FdtdMatCylinder);
    if (need > s->mat_stage_len) {
        (void)hipFree(s->mat_stage);
        s->mat_stage = nullptr;
        s->mat_stage_len = 0;
        HIP_TRY(hipMalloc(&s->mat_stage, need));
        s->mat_stage_len = need;
    }
    int32_t *d_surf = m.nlayers > 0 ? (int32_t *)s->mat_stage : nullptr;
    FdtdMatCylinder *d_cyl = m.ncyl > 0 ? (FdtdMatCylinder *)((char *)s->mat_stage + surf_bytes) : nullptr;
    float *Ca = s->mat, *Cb = s->mat + s->ncell;
    int rc = FDTD_OK;
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == FDTD_OK) rc = fail(FDTD_EHIP, "%s: %s", what, hipGetErrorString(e));
        return rc == FDTD_OK;
    };
    if (m.nlayers > 0)
        hip_ok(hipMemcpyAsync(d_surf, surfaces, m.nlayers * plane * sizeof(int32_t), hipMemcpyHostToDevice,
                              s->self.stream), "hipMemcpyAsync");
    if (rc == FDTD_OK && m.ncyl > 0 &&
        hip_ok(hipMemcpyAsync(d_cyl, cyl, m.ncyl * sizeof(FdtdMatCylinder), hipMemcpyHostToDevice,
                              s->self.stream), "hipMemcpyAsync") &&
        hip_ok(hipMemsetAsync(Cb, 0, s->ncell * sizeof(float), s->self.stream), "hipMemsetAsync")) {
        // owners of the cylinder cells, in Cb's storage
        hipLaunchKernelGGL(k_mat_cylinders, dim3(m.ncyl, 64), dim3(256), 0, s->self.stream, d_cyl,
                           (double)s->c.cell_size, reinterpret_cast<int *>(Cb), (long long)s->pitch,
                           (long long)s->pitch * ny);
        hip_ok(hipGetLastError(), "k_mat_cylinders");
    }
    if (rc == FDTD_OK) {
        const long long n = (long long)plane * nz;
        hipLaunchKernelGGL(k_mat_build, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->self.stream, m,
                           d_surf, d_cyl, nx, ny, nz, (long long)s->pitch, (long long)s->pitch * ny, Ca, Cb);
        hip_ok(hipGetLastError(), "k_mat_build");
    }
    hip_ok(hipStreamSynchronize(s->self.stream), "hipStreamSynchronize");

    if (rc) return rc;
    // planes from nz_mat up are the vacuum compute_material_arrays appends;
    // its Ca, Cb by the same float32 operations
    const float perm = 1.0f * m.eps0, loss = (0.0f * m.dt) / (2.0f * perm);
    s->h.uni_z = m.nz_mat;
    s->h.ca_u = (1.0f - loss) / (1.0f + loss);
    s->h.cb_u = (m.dt / (perm * m.dx)) / (1.0f + loss);
    s->mat_set = true;
    return sim_refresh(s);
}

extern "C" int fdtd_sim_set_uniform_material(FdtdSim *s, float ca, float cb) {
    CHECK(s, "NULL sim");
    s->h.uni_z = 0;
    s->h.ca_u = ca;
    s->h.cb_u = cb;
    s->mat_set = true;
    return sim_refresh(s);
}

extern "C" int fdtd_sim_set_source(FdtdSim *s, const float *src_ex, const float *src_ey, int n_src) {
    CHECK(s && src_ex && src_ey, "NULL argument");
    CHECK(s->g.mode == FDTD_MODE_INCIDENT, "hard sources are for INCIDENT simulations");
    CHECK(n_src >= 0, "negative n_src");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    (void)hipFree(s->src);
    s->src = nullptr;
    const size_t plane = (size_t)s->g.ny * s->g.nx;
    const size_t n = (size_t)n_src * plane;
    if (n) {
        HIP_TRY(hipMalloc(&s->src, 2 * n * sizeof(float)));
        HIP_TRY(hipMemcpy(s->src, src_ex, n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->src + n, src_ey, n * sizeof(float), hipMemcpyHostToDevice));
    }
    s->h.src = s->src;
    s->h.src_cs = (long long)n;
    s->h.n_src = n_src;
    return sim_refresh(s);
}

extern "C" int fdtd_sim_attach_beam(FdtdSim *s, FdtdBeam *b) {
    CHECK(s && b, "NULL argument");
    CHECK(b->ny == s->g.ny && b->nx == s->g.nx, "beam plane %d x %d does not match grid %d x %d", b->ny,
          b->nx, s->g.ny, s->g.nx);
    CHECK(b->device == s->device, "beam and simulation are on different devices");
    if (s->g.mode == FDTD_MODE_MAIN)
        CHECK(b->nt >= s->g.nt_inc, "beam has %d steps, simulation injects %d", b->nt, s->g.nt_inc);
    else
        CHECK(b->nt >= s->g.nt, "beam has %d steps, incident run records %d", b->nt, s->g.nt);
    b->refs++;
    beam_unref(s->beam);
    s->beam = b;
    s->h.nt_beam = b->nt;
    s->h.beam_cs = (long long)b->nt * b->ny * b->nx;
    if (s->g.mode == FDTD_MODE_MAIN) {
        s->h.beam = b->d;
        s->h.rec = nullptr;
        s->h.dft_items = s->h.face_off[s->h.nfaces];
    } else {
        s->h.beam = nullptr;
        s->h.rec = b->d;
        s->h.dft_items = s->h.face_off[s->h.nfaces] + 4LL * b->ny * b->nx;
    }
    return sim_refresh(s);
}

extern "C" int fdtd_sim_reset(FdtdSim *s) {
    CHECK(s, "NULL sim");
    HIP_TRY(hipSetDevice(s->device));
    const Sizes z = sizes_of(&s->g);
    hipStream_t st = s->self.stream;
    const int ns = nsets_of(&s->g);
    HIP_TRY(hipMemsetAsync(s->fields, 0, 6 * field_stride(z.ncell) * ns * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(s->psi, 0, (z.psi_total + (ns - 1) * psi_h_len(z)) * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(s->acc, 0, z.acc_total * sizeof(float4), st));
    HIP_TRY(hipMemsetAsync(s->step, 0, 2 * sizeof(int), st));
    HIP_TRY(hipStreamSynchronize(st));
    return FDTD_OK;
}

static int check_ready(FdtdSim *s) {
    if (s->g.mode == FDTD_MODE_MAIN) {
        CHECK(s->mat_set, "materials not set");
        if (s->g.nt_inc > 0) CHECK(s->beam, "no beam attached (simulation injects %d steps)", s->g.nt_inc);
    } else {
        CHECK(s->mat_set && s->h.uni_z == 0, "INCIDENT simulations need a uniform material");
        CHECK(s->g.nt_inc == 0 || s->src, "hard source not set");
    }
    return FDTD_OK;
}

extern "C" int fdtd_sim_run(FdtdSim *s, int nsteps) {
    CHECK(s, "NULL sim");
    int rc = check_ready(s);
    if (rc) return rc;
    return group_run(s->self, nsteps, false);
}

extern "C" int fdtd_sim_sync(FdtdSim *s) {
    CHECK(s, "NULL sim");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    return FDTD_OK;
}

extern "C" int fdtd_sim_step_count(FdtdSim *s, int *t) {
    CHECK(s && t, "NULL argument");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(t, s->step, sizeof(int), hipMemcpyDeviceToHost, s->self.stream));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    return FDTD_OK;
}

static int ensure_scratch(FdtdSim *s, long long n) {
    if (s->scratch_len >= n) return FDTD_OK;
    (void)hipFree(s->scratch_c64);
    s->scratch_c64 = nullptr;
    s->scratch_len = 0;
    HIP_TRY(hipMalloc(&s->scratch_c64, n * sizeof(float2)));
    s->scratch_len = n;
    return FDTD_OK;
}

extern "C" int fdtd_sim_get_phasor(FdtdSim *s, int index, float *out_c64) {
    CHECK(s && out_c64, "NULL argument");
    CHECK(index >= 0 && index < 4 * s->h.nfaces, "phasor index %d outside [0, %d)", index, 4 * s->h.nfaces);
    HIP_TRY(hipSetDevice(s->device));
    const int f = index / 4;
    const long long n = s->h.face_off[f + 1] - s->h.face_off[f];
    int rc = ensure_scratch(s, n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_phasor_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->self.stream,
                       s->acc + s->h.acc_off[index], n, s->scratch_c64);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_c64, s->scratch_c64, n * sizeof(float2), hipMemcpyDeviceToHost, s->self.stream));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    return FDTD_OK;
}

// Buffer set holding the current fields (reads the device step counter).
static int host_current_set(FdtdSim *s, int *set) {
    int t = 0;
    int rc = fdtd_sim_step_count(s, &t);
    if (rc) return rc;
    *set = current_set(s->h, t);
    return FDTD_OK;
}

extern "C" int fdtd_sim_get_materials(FdtdSim *s, float *Ca, float *Cb) {
    CHECK(s && Ca && Cb, "NULL argument");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "per-cell materials are for MAIN simulations");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    int rc = copy_volume_out(s, Ca, s->mat);
    if (!rc) rc = copy_volume_out(s, Cb, s->mat + s->ncell);
    return rc;
}

extern "C" int fdtd_sim_get_field(FdtdSim *s, int comp, float *out) {
    CHECK(s && out, "NULL argument");
    CHECK(comp >= 0 && comp < 6, "field component %d outside [0, 6)", comp);
    HIP_TRY(hipSetDevice(s->device));
    int set = 0;
    int rc = host_current_set(s, &set);
    if (rc) return rc;
    return copy_volume_out(s, out, s->h.f[set][comp]);
}

extern "C" int fdtd_sim_set_field(FdtdSim *s, int comp, const float *in) {
    CHECK(s && in, "NULL argument");
    CHECK(comp >= 0 && comp < 6, "field component %d outside [0, 6)", comp);
    HIP_TRY(hipSetDevice(s->device));
    int set = 0;
    int rc = host_current_set(s, &set);
    if (rc) return rc;
    rc = copy_volume_in(s, s->h.f[set][comp], in);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->self.stream));
    return FDTD_OK;
}

extern "C" int fdtd_sim_get_psi(FdtdSim *s, int index, float *out) {
    CHECK(s && out, "NULL argument");
    CHECK(index >= 0 && index < 24, "psi index %d outside [0, 24)", index);
    HIP_TRY(hipSetDevice(s->device));
    const int face = (index % 12) / 2;
    const FdtdGeom &g = s->g;
    int set = 0;
    int rc = host_current_set(s, &set);
    if (rc) return rc;
    const float *src = s->h.psi[set][index];
    if (face == 2 || face == 3) {   // l/r: compact (nz, ny, npml)
        HIP_TRY(hipMemcpyAsync(out, src, (size_t)s->psi_len[index] * sizeof(float), hipMemcpyDeviceToHost,
                               s->self.stream));
        HIP_TRY(hipStreamSynchronize(s->self.stream));
        return FDTD_OK;
    }
    const size_t rows = (face < 2) ? (size_t)g.npml * g.ny : (size_t)g.nz * g.npml;
    return copy_rows_out(s, out, src, rows);
}

extern "C" int fdtd_sim_set_stepper(FdtdSim *s, int fused) {
    CHECK(s, "NULL sim");
    CHECK(fused >= 0 && fused <= 3,
          "stepper must be 0 (split), 1 (fused strips), 2 (fused rows) or 3 (fused, halo recomputed)");
    CHECK(!fused || s->g.mode == FDTD_MODE_MAIN, "the fused stepper runs MAIN simulations");
    int t = 0;
    int rc = fdtd_sim_step_count(s, &t);
    if (rc) return rc;
    CHECK(t == 0, "the stepper can only change before the first step (reset first)");
    if (s->h.fused == fused) return FDTD_OK;
    s->h.fused = fused;
    s->self.launch = default_launch(fused);
    return sim_refresh(s);
}

extern "C" int fdtd_sim_set_launch(FdtdSim *s, int ty, int zchunk, int xvec, int xwaves, int xcd_order) {
    CHECK(s, "NULL sim");
    CHECK(zchunk >= 0 && zchunk < 0x8000, "zchunk out of range");
    const Launch keep = s->self.launch;
    s->self.launch.ty = ty;
    s->self.launch.zchunk = zchunk;
    s->self.launch.xvec = xvec;
    s->self.launch.xw = xwaves;
    s->self.launch.xcd = xcd_order;
    int rc = group_build(s->self);
    if (rc) {
        std::string err = g_err;
        s->self.launch = keep;
        (void)group_build(s->self);
        g_err = err;
    }
    return rc;
}

static int time_group(Group &G, int nsteps, bool graph, float *ms_total, float *ms_h, float *ms_e,
                      float *ms_dft) {
    CHECK(nsteps > 0, "nsteps must be positive");
    HIP_TRY(hipSetDevice(G.device));
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    HIP_TRY(hipEventRecord(ev[0], G.stream));
    int rc = group_run(G, nsteps, graph);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev[1], G.stream));
    HIP_TRY(hipEventSynchronize(ev[1]));
    float tot = 0;
    HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev[1]));
    if (ms_total) *ms_total = tot;
    if (ms_h || ms_e || ms_dft) {
        // per-kernel: bracket each launch of the same step sequence
        const KernelId order_main[3] = {KH, KE, KD}, order_inc[3] = {KE, KH, KD}, order_fused[2] = {KF, KD};
        const KernelId *order = G.mode != FDTD_MODE_MAIN ? order_inc : G.fused ? order_fused : order_main;
        const int nk = (G.mode == FDTD_MODE_MAIN && G.fused) ? 2 : 3;
        std::vector<hipEvent_t> evs(2 * nk * (size_t)nsteps);
        for (auto &e : evs) HIP_TRY(hipEventCreate(&e));
        for (int s = 0; s < nsteps; ++s)
            for (int k = 0; k < nk; ++k) {
                HIP_TRY(hipEventRecord(evs[2 * nk * s + 2 * k], G.stream));
                launch_kernel(G, order[k]);
                HIP_TRY(hipEventRecord(evs[2 * nk * s + 2 * k + 1], G.stream));
            }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(evs.back()));
        double acc[4] = {0, 0, 0, 0};
        for (int s = 0; s < nsteps; ++s)
            for (int k = 0; k < nk; ++k) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, evs[2 * nk * s + 2 * k], evs[2 * nk * s + 2 * k + 1]));
                acc[order[k]] += ms;
            }
        for (auto &e : evs) (void)hipEventDestroy(e);
        // fused stepper: ms_h reports the fused H+E kernel, ms_e is 0
        if (ms_h) *ms_h = (float)((acc[KH] + acc[KF]) / nsteps);
        if (ms_e) *ms_e = (float)(acc[KE] / nsteps);
        if (ms_dft) *ms_dft = (float)(acc[KD] / nsteps);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return FDTD_OK;
}

extern "C" int fdtd_sim_time_steps(FdtdSim *s, int nsteps, float *ms_total, float *ms_h, float *ms_e,
                                   float *ms_dft) {
    CHECK(s, "NULL sim");
    int rc = check_ready(s);
    if (rc) return rc;
    return time_group(s->self, nsteps, false, ms_total, ms_h, ms_e, ms_dft);
}

// ---------------------------------------------------------------------------
// near-to-far
// ---------------------------------------------------------------------------
// Device work area of one near-to-far call, carved from one buffer in
// 256-byte aligned pieces: the direction lists (theta, phi), the column and
// row phase tables, the top-face factors, the per-(direction, split)
// partial sums and the output.
struct N2fWork {
    double *thetas, *phis;
    double2 *phx, *phy, *zph;
    double4 *trig;
    double *part;
    double2 *out;
};

static constexpr size_t n2f_round(size_t b) { return (b + 255) & ~(size_t)255; }

static int n2f_splits(int nyp) { return std::max(1, std::min(64, nyp / 8)); }

static size_t n2f_carve(char *base, int ndir, int nxp, int nyp, N2fWork *w) {
    const size_t T = (size_t)ndir;
    const size_t sizes[8] = {T * sizeof(double), T * sizeof(double), T * nxp * sizeof(double2),
                             T * nyp * sizeof(double2), T * sizeof(double2), T * sizeof(double4),
                             T * n2f_splits(nyp) * 8 * sizeof(double), 4 * T * sizeof(double2)};
    void *ptrs[8];
    size_t off = 0;
    for (int i = 0; i < 8; ++i) {
        ptrs[i] = base ? base + off : nullptr;
        off += n2f_round(sizes[i]);
    }
    if (w) {
        w->thetas = (double *)ptrs[0];
        w->phis = (double *)ptrs[1];
        w->phx = (double2 *)ptrs[2];
        w->phy = (double2 *)ptrs[3];
        w->zph = (double2 *)ptrs[4];
        w->trig = (double4 *)ptrs[5];
        w->part = (double *)ptrs[6];
        w->out = (double2 *)ptrs[7];
    }
    return off;
}

// Preallocated work area of fdtd_sim_near2far_noalloc: one per device (each
// GPU the library runs on holds its own copy until the process ends), sized
// for FDTD_N2F_MAX_THETA angles and FDTD_N2F_MAX_SIDE samples per side.
// Host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (n2f_scratch_ptr), never through its name.
static constexpr size_t kN2fScratchBytes =
    2 * n2f_round(FDTD_N2F_MAX_THETA * sizeof(double)) +
    2 * n2f_round((size_t)FDTD_N2F_MAX_THETA * FDTD_N2F_MAX_SIDE * sizeof(double2)) +
    n2f_round(FDTD_N2F_MAX_THETA * sizeof(double2)) + n2f_round(FDTD_N2F_MAX_THETA * sizeof(double4)) +
    n2f_round((size_t)FDTD_N2F_MAX_THETA * 64 * 8 * sizeof(double)) +
    n2f_round(4 * FDTD_N2F_MAX_THETA * sizeof(double2));
static_assert(kN2fScratchBytes == FDTD_N2F_SCRATCH_BYTES, "header states the scratch size");
static __device__ __align__(256) char scratch[kN2fScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on `device`, looked up once per device with
// that device current.
static int n2f_scratch_ptr(int device, void **out) {
    CHECK(device >= 0 && device < kMaxDevices, "device %d beyond the library's %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        HIP_TRY(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(FDTD_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return FDTD_OK;
}

// Near-to-far transform on stream st for ndir directions (thetas_deg[i],
// phis_deg[i]).  work: a device area of at least n2f_carve(nullptr, ...)
// bytes, or NULL to allocate one for the call.  Every piece of the area is
// written before it is read, so nothing needs clearing between calls.
// out: 4 x ndir complex128 (E_theta, E_phi, H_theta, H_phi).
template <typename P>
static int near2far_run(int device, hipStream_t st, const P *ex, const P *ey, const P *hx, const P *hy,
                        int nxp, int nyp, int nzp, double dx, const double *thetas_deg,
                        const double *phis_deg, int ndir, double r, double freq, double *out, void *work) {
    CHECK(nxp >= 3 && nyp >= 3, "phasor plane %d x %d too small", nyp, nxp);
    CHECK(nxp % 2 == 1 && nyp % 2 == 1, "composite Simpson needs odd sample counts (got %d x %d)", nyp, nxp);
    CHECK(ndir > 0 && thetas_deg && phis_deg && out, "bad angle arguments");
    const double pi = 3.141592653589793;
    const double eta0 = 376.730313668, c0 = 299792458.0;
    const double lam = c0 / freq, k = 2.0 * pi / lam;
    const double z_top = (double)((nzp - 1) / 2) * dx;   // z_loc[-1] (near2far.py:73, :97)
    // ff_weight = 1j * k * exp(-1j * k * r) / (4 pi r)   (near2far.py:76-77)
    const double kr = k * r;
    const double2 wgt = make_double2(k * std::sin(kr) / (4.0 * pi * r), k * std::cos(kr) / (4.0 * pi * r));
    const int nsplit = n2f_splits(nyp);
    HIP_TRY(hipSetDevice(device));
    char *own = nullptr;
    if (!work) {
        HIP_TRY(hipMalloc(&own, n2f_carve(nullptr, ndir, nxp, nyp, nullptr)));
        work = own;
    }
    N2fWork w;
    n2f_carve((char *)work, ndir, nxp, nyp, &w);
    int rc = FDTD_OK;
    hipError_t e = hipMemcpyAsync(w.thetas, thetas_deg, (size_t)ndir * sizeof(double),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w.phis, phis_deg, (size_t)ndir * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_n2f_tables, dim3(ndir, (std::max(nxp, nyp) + 255) / 256), dim3(256), 0, st,
                           w.thetas, w.phis, nxp, nyp, dx, k, z_top, w.phx, w.phy, w.zph, w.trig);
        hipLaunchKernelGGL(k_n2f_partial<P>, dim3(ndir, nsplit), dim3(256), 0, st, ex, ey, hx, hy, nxp, nyp,
                           nsplit, w.phx, w.phy, w.part);
        hipLaunchKernelGGL(k_n2f_final, dim3((ndir + 63) / 64), dim3(64), 0, st, w.part, ndir, nsplit,
                           w.trig, w.zph, wgt, eta0, w.out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, w.out, (size_t)4 * ndir * sizeof(double2), hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) rc = fail(FDTD_EHIP, "near-to-far on device %d failed: %s", device, hipGetErrorString(e));
    if (own) (void)hipFree(own);
    return rc;
}

// Direction lists of a theta x phi grid, phi-major (the reference's
// (nphi, ntheta) output order).
static void n2f_grid(const double *thetas, int ntheta, const double *phis, int nphi, std::vector<double> &th,
                     std::vector<double> &ph) {
    th.resize((size_t)ntheta * nphi);
    ph.resize(th.size());
    for (int p = 0; p < nphi; ++p)
        for (int t = 0; t < ntheta; ++t) {
            th[(size_t)p * ntheta + t] = thetas[t];
            ph[(size_t)p * ntheta + t] = phis[p];
        }
}

static int near2far_of_sim(FdtdSim *s, const double *thetas_deg, const double *phis_deg, int ndir, double r,
                           double freq, double *out, void *work) {
    HIP_TRY(hipSetDevice(s->device));
    const float4 *u = s->acc;   // ex_ph_u ey_ph_u hx_ph_u hy_ph_u = phasors 4..7
    const int np = s->g.npml;
    return near2far_run<float4>(s->device, s->self.stream, u + s->h.acc_off[4], u + s->h.acc_off[5],
                                u + s->h.acc_off[6], u + s->h.acc_off[7], s->g.nx - 2 * np, s->g.ny - 2 * np,
                                s->g.nz - 2 * np, s->c.cell_size, thetas_deg, phis_deg, ndir, r, freq, out,
                                work);
}

extern "C" int fdtd_sim_near2far(FdtdSim *s, const double *thetas_deg, int ntheta, double phi_deg, double r,
                                 double freq, double *out) {
    CHECK(s, "NULL sim");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "near-to-far needs a MAIN simulation");
    CHECK(ntheta > 0 && thetas_deg, "bad theta arguments");
    std::vector<double> th, ph;
    n2f_grid(thetas_deg, ntheta, &phi_deg, 1, th, ph);
    return near2far_of_sim(s, th.data(), ph.data(), ntheta, r, freq, out, nullptr);
}

extern "C" int fdtd_sim_near2far_grid(FdtdSim *s, const double *thetas_deg, int ntheta, const double *phis_deg,
                                      int nphi, double r, double freq, double *out) {
    CHECK(s, "NULL sim");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "near-to-far needs a MAIN simulation");
    CHECK(ntheta > 0 && nphi > 0 && thetas_deg && phis_deg, "bad angle arguments");
    CHECK((long long)ntheta * nphi < (1LL << 31), "too many directions");
    std::vector<double> th, ph;
    n2f_grid(thetas_deg, ntheta, phis_deg, nphi, th, ph);
    return near2far_of_sim(s, th.data(), ph.data(), (int)th.size(), r, freq, out, nullptr);
}

extern "C" int fdtd_sim_near2far_noalloc(FdtdSim *s, const double *thetas_deg, int ntheta, double phi_deg,
                                         double r, double freq, double *out) {
    CHECK(s, "NULL sim");
    CHECK(s->g.mode == FDTD_MODE_MAIN, "near-to-far needs a MAIN simulation");
    CHECK(ntheta >= 1 && ntheta <= FDTD_N2F_MAX_THETA, "ntheta %d outside [1, %d] (preallocated work area)",
          ntheta, FDTD_N2F_MAX_THETA);
    CHECK(thetas_deg, "bad theta arguments");
    const int nxp = s->g.nx - 2 * s->g.npml, nyp = s->g.ny - 2 * s->g.npml;
    CHECK(nxp <= FDTD_N2F_MAX_SIDE && nyp <= FDTD_N2F_MAX_SIDE,
          "phasor plane %d x %d exceeds %d samples per side (preallocated work area)", nyp, nxp,
          FDTD_N2F_MAX_SIDE);
    void *work = nullptr;
    int rc = n2f_scratch_ptr(s->device, &work);
    if (rc) return rc;
    // the direction lists live on the stack: no allocation
    double th[FDTD_N2F_MAX_THETA], ph[FDTD_N2F_MAX_THETA];
    for (int t = 0; t < ntheta; ++t) {
        th[t] = thetas_deg[t];
        ph[t] = phi_deg;
    }
    return near2far_of_sim(s, th, ph, ntheta, r, freq, out, work);
}

static int near2far_host_planes(const float *ex_u, const float *ey_u, const float *hx_u, const float *hy_u,
                                int nx_ph, int ny_ph, int nz_ph, double dx, const double *th, const double *ph,
                                int ndir, double r, double freq, double *out) {
    CHECK(ex_u && ey_u && hx_u && hy_u, "NULL phasor");
    CHECK(nx_ph > 0 && ny_ph > 0, "bad plane size");
    HIP_TRY(hipSetDevice(g_device));
    const size_t n = (size_t)nx_ph * ny_ph;
    float2 *d;
    HIP_TRY(hipMalloc(&d, 4 * n * sizeof(float2)));
    hipStream_t st;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const float *src[4] = {ex_u, ey_u, hx_u, hy_u};
    for (int c = 0; c < 4; ++c)
        HIP_TRY(hipMemcpyAsync(d + c * n, src[c], n * sizeof(float2), hipMemcpyHostToDevice, st));
    int rc = near2far_run<float2>(g_device, st, d, d + n, d + 2 * n, d + 3 * n, nx_ph, ny_ph, nz_ph, dx, th, ph,
                                  ndir, r, freq, out, nullptr);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    (void)hipFree(d);
    return rc;
}

extern "C" int fdtd_near2far_host(const float *ex_u, const float *ey_u, const float *hx_u, const float *hy_u,
                                  int nx_ph, int ny_ph, int nz_ph, double dx, const double *thetas_deg,
                                  int ntheta, double phi_deg, double r, double freq, double *out) {
    CHECK(ntheta > 0 && thetas_deg, "bad theta arguments");
    std::vector<double> th, ph;
    n2f_grid(thetas_deg, ntheta, &phi_deg, 1, th, ph);
    return near2far_host_planes(ex_u, ey_u, hx_u, hy_u, nx_ph, ny_ph, nz_ph, dx, th.data(), ph.data(), ntheta,
                                r, freq, out);
}

extern "C" int fdtd_near2far_host_grid(const float *ex_u, const float *ey_u, const float *hx_u,
                                       const float *hy_u, int nx_ph, int ny_ph, int nz_ph, double dx,
                                       const double *thetas_deg, int ntheta, const double *phis_deg, int nphi,
                                       double r, double freq, double *out) {
    CHECK(ntheta > 0 && nphi > 0 && thetas_deg && phis_deg, "bad angle arguments");
    CHECK((long long)ntheta * nphi < (1LL << 31), "too many directions");
    std::vector<double> th, ph;
    n2f_grid(thetas_deg, ntheta, phis_deg, nphi, th, ph);
    return near2far_host_planes(ex_u, ey_u, hx_u, hy_u, nx_ph, ny_ph, nz_ph, dx, th.data(), ph.data(),
                                (int)th.size(), r, freq, out);
}

// ---------------------------------------------------------------------------
// beams
// ---------------------------------------------------------------------------
extern "C" int fdtd_beam_create(int nt, int ny, int nx, const float *ex, const float *ey, const float *hx,
                                const float *hy, FdtdBeam **out) {
    CHECK(out, "out is NULL");
    CHECK(nt >= 0 && ny > 0 && nx > 0, "bad beam shape (%d, %d, %d)", nt, ny, nx);
    const bool have = ex || ey || hx || hy;
    CHECK(!have || (ex && ey && hx && hy), "give all four beam components or none");
    HIP_TRY(hipSetDevice(g_device));
    FdtdBeam *b = new FdtdBeam();
    b->device = g_device;
    b->nt = nt; b->ny = ny; b->nx = nx;
    const size_t n = (size_t)nt * ny * nx;
    hipError_t e = hipMalloc(&b->d, std::max<size_t>(1, 4 * n) * sizeof(float));
    if (e != hipSuccess) {
        delete b;
        return fail(FDTD_ENOMEM, "beam allocation of %zu bytes failed: %s", 4 * n * sizeof(float),
                    hipGetErrorString(e));
    }
    const float *src[4] = {ex, ey, hx, hy};
    for (int c = 0; c < 4; ++c) {
        e = have ? hipMemcpy(b->d + c * n, src[c], n * sizeof(float), hipMemcpyHostToDevice)
                 : hipMemset(b->d + c * n, 0, n * sizeof(float));
        if (e != hipSuccess) {
            (void)hipFree(b->d);
            delete b;
            return fail(FDTD_EHIP, "beam upload failed: %s", hipGetErrorString(e));
        }
    }
    *out = b;
    return FDTD_OK;
}

extern "C" int fdtd_beam_upload(FdtdBeam *b, int t0, int n, const float *ex, const float *ey, const float *hx,
                                const float *hy) {
    CHECK(b && ex && ey && hx && hy, "NULL argument");
    CHECK(t0 >= 0 && n >= 0 && t0 + n <= b->nt, "steps [%d, %d) outside the beam's %d", t0, t0 + n, b->nt);
    HIP_TRY(hipSetDevice(b->device));
    const size_t plane = (size_t)b->ny * b->nx, total = (size_t)b->nt * plane;
    const float *src[4] = {ex, ey, hx, hy};
    for (int c = 0; c < 4; ++c)
        HIP_TRY(hipMemcpy(b->d + c * total + (size_t)t0 * plane, src[c], (size_t)n * plane * sizeof(float),
                          hipMemcpyHostToDevice));
    return FDTD_OK;
}

extern "C" int fdtd_beam_get(FdtdBeam *b, int comp, float *out) {
    CHECK(b && out, "NULL argument");
    CHECK(comp >= 0 && comp < 4, "beam component %d outside [0, 4)", comp);
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipDeviceSynchronize());
    const size_t n = (size_t)b->nt * b->ny * b->nx;
    HIP_TRY(hipMemcpy(out, b->d + comp * n, n * sizeof(float), hipMemcpyDeviceToHost));
    return FDTD_OK;
}

extern "C" void fdtd_beam_release(FdtdBeam *b) { beam_unref(b); }

int beam_synthesize_source(const double *pulse_fft, const double *all_k, int nt_inc, int n_src, int nx, int ny,
                           double dx, double theta_deg, double phi_deg, double radius, int pol, hipStream_t st,
                           float **src_out, char *err, size_t errlen);

extern "C" int fdtd_beam_synthesize(const FdtdGeom *g, const FdtdCoeffs *c, const FdtdBeamSpec *spec,
                                    FdtdBeam **out, float *inc_phasors_c64) {
    CHECK(g && c && spec && out && inc_phasors_c64, "NULL argument");
    CHECK(g->mode == FDTD_MODE_INCIDENT, "beam synthesis runs an INCIDENT simulation");
    CHECK(spec->pulse_fft && spec->all_k, "pulse spectrum not given");
    CHECK(g->nt > 0 && g->nt_inc >= 0 && g->nt_inc <= g->nt, "bad step counts (nt %d, n_src %d)", g->nt,
          g->nt_inc);
    *out = nullptr;
    FdtdSim *s = nullptr;
    int rc = fdtd_sim_create(g, c, &s);
    if (rc) return rc;
    FdtdBeam *b = nullptr;
    auto cleanup = [&](int code) {
        std::string keep = g_err;
        fdtd_sim_destroy(s);
        if (b) beam_unref(b);
        g_err = keep;
        return code;
    };
    rc = fdtd_sim_set_uniform_material(s, 1.0f, spec->cb_vacuum);
    if (rc) return cleanup(rc);
    float *d_src = nullptr;
    char err[256] = {0};
    if (beam_synthesize_source(spec->pulse_fft, spec->all_k, g->nt, g->nt_inc, g->nx, g->ny, c->cell_size,
                               spec->theta_deg, spec->phi_deg, spec->gauss_radius, spec->polarization,
                               s->self.stream, &d_src, err, sizeof err) != 0)
        return cleanup(fail(FDTD_EHIP, "beam synthesis: %s", err));
    s->src = d_src;                                 // the sim owns it now
    s->h.src = d_src;
    s->h.src_cs = (long long)g->nt_inc * g->ny * g->nx;
    s->h.n_src = g->nt_inc;
    rc = sim_refresh(s);
    if (!rc) rc = fdtd_beam_create(g->nt, g->ny, g->nx, nullptr, nullptr, nullptr, nullptr, &b);
    if (!rc) rc = fdtd_sim_attach_beam(s, b);
    if (!rc) rc = fdtd_sim_run(s, g->nt);
    if (!rc) rc = fdtd_sim_sync(s);
    const size_t n = (size_t)(g->ny - 2 * g->npml) * (g->nx - 2 * g->npml);
    for (int i = 0; i < 4 && !rc; ++i) rc = fdtd_sim_get_phasor(s, i, inc_phasors_c64 + 2 * n * i);
    if (rc) return cleanup(rc);
    fdtd_sim_destroy(s);
    *out = b;
    return FDTD_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static void batch_clear_graphs(FdtdBatch *b) {
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : b->graph_src) (void)hipGraphDestroy(kv.second);
    b->graphs.clear();
    b->graph_src.clear();
}

static long long sim_cells(const FdtdSim *s) { return (long long)s->g.nx * s->g.ny * s->g.nz; }

// Split the members between the packed group and the serial list for the
// batch's mode and deal the serial members to lanes, then rebuild the
// packed group's tile list.  Lanes: in SERIAL mode with small_cells > 0 the
// members below small_cells are dealt over the lanes (largest first, each
// to the lane with the fewest cells so far) and step first, concurrently;
// the others then step one after another on the batch stream (lane -1).
// Otherwise every serial member is dealt over the lanes.
static int batch_build(FdtdBatch *b) {
    batch_clear_graphs(b);
    std::vector<FdtdSim *> packed;
    b->serial.clear();
    for (FdtdSim *s : b->all) {
        const bool pack = b->mode == FDTD_BATCH_PACKED ||
                          (b->mode == FDTD_BATCH_HYBRID && sim_cells(s) < b->small_cells);
        (pack ? packed : b->serial).push_back(s);
    }
    std::stable_sort(b->serial.begin(), b->serial.end(),
                     [](const FdtdSim *a, const FdtdSim *c) { return sim_cells(a) > sim_cells(c); });
    const bool split = b->mode == FDTD_BATCH_SERIAL && b->small_cells > 0;
    std::vector<long long> load(b->nlanes, 0);
    b->lane.assign(b->serial.size(), 0);
    for (size_t k = 0; k < b->serial.size(); ++k) {
        if (split && sim_cells(b->serial[k]) >= b->small_cells) {
            b->lane[k] = -1;
            continue;
        }
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->lane[k] = l;
        load[l] += sim_cells(b->serial[k]);
    }
    while ((int)b->streams.size() < b->nlanes) {
        hipStream_t st;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->streams.push_back(st);
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->fork.push_back(ev);
    }
    b->grp.members = packed;
    if (packed.empty()) {
        group_clear_graphs(b->grp);
        b->grp.ntiles = b->grp.nmodels = 0;
        return FDTD_OK;
    }
    return group_build(b->grp);
}

// The launches of len steps: the packed group step by step, then the
// members dealt to lanes, each member's len steps with its own tile shape
// on its lane's stream, then the members of lane -1 one after another on
// the batch stream.  Lanes other than the batch stream start after an event
// recorded on it and end with an event it waits for, so a chunk is one
// fork-join (and one graph when captured).  Members are independent, so
// neither the order nor the concurrency changes results; a member's steps
// back to back keep its working set in the caches.
static int batch_launch(FdtdBatch *b, int len) {
    hipStream_t st0 = b->grp.stream;
    bool dealt = false;
    for (int l : b->lane) dealt = dealt || l > 0;
    if (dealt) {
        HIP_TRY(hipEventRecord(b->fork[0], st0));
        for (int l = 1; l < b->nlanes; ++l) HIP_TRY(hipStreamWaitEvent(b->streams[l], b->fork[0], 0));
    }
    if (!b->grp.members.empty())
        for (int s = 0; s < len; ++s) launch_step(b->grp);
    for (size_t k = 0; k < b->serial.size(); ++k) {
        if (b->lane[k] < 0) continue;
        hipStream_t st = b->lane[k] == 0 ? st0 : b->streams[b->lane[k]];
        for (int s = 0; s < len; ++s) launch_step(b->serial[k]->self, st);
    }
    if (dealt)
        for (int l = 1; l < b->nlanes; ++l) {
            HIP_TRY(hipEventRecord(b->fork[l], b->streams[l]));
            HIP_TRY(hipStreamWaitEvent(st0, b->fork[l], 0));
        }
    for (size_t k = 0; k < b->serial.size(); ++k)
        if (b->lane[k] < 0)
            for (int s = 0; s < len; ++s) launch_step(b->serial[k]->self, st0);
    return FDTD_OK;
}

static int batch_run(FdtdBatch *b, int nsteps, bool use_graph) {
    CHECK(nsteps >= 0, "negative step count");
    if (b->mode == FDTD_BATCH_PACKED) return group_run(b->grp, nsteps, use_graph);
    if (nsteps == 0) return FDTD_OK;
    HIP_TRY(hipSetDevice(b->grp.device));
    hipStream_t st = b->grp.stream;
    if (!use_graph) {
        int rc = batch_launch(b, nsteps);
        if (rc) return rc;
        HIP_TRY(hipGetLastError());
        return FDTD_OK;
    }
    const int chunk = std::min(nsteps, 64);
    auto replay = [&](int len) -> int {
        auto it = b->graphs.find(len);
        if (it == b->graphs.end()) {
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = batch_launch(b, len);
            hipError_t ec = hipStreamEndCapture(st, &graph);
            if (rc) return rc;
            HIP_TRY(ec);
            hipGraphExec_t exec;
            HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            b->graph_src[len] = graph;
            it = b->graphs.emplace(len, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, st));
        return FDTD_OK;
    };
    int done = 0;
    while (done + chunk <= nsteps) {
        int rc = replay(chunk);
        if (rc) return rc;
        done += chunk;
    }
    if (done < nsteps) return replay(nsteps - done);
    return FDTD_OK;
}

extern "C" int fdtd_batch_create(FdtdSim *const *sims, int n, FdtdBatch **out) {
    CHECK(sims && n > 0 && out, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        CHECK(sims[i], "sim %d is NULL", i);
        int rc = check_ready(sims[i]);
        if (rc) return rc;
        CHECK(sims[i]->device == sims[0]->device, "batch members must share a device");
    }
    FdtdBatch *b = new FdtdBatch();
    b->grp.device = sims[0]->device;
    b->all.assign(sims, sims + n);
    b->grp.launch = default_launch(sims[0]->h.fused);
    (void)hipSetDevice(b->grp.device);
    hipError_t e = hipStreamCreateWithFlags(&b->grp.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(FDTD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    b->streams.push_back(b->grp.stream);
    hipEvent_t ev0;
    e = hipEventCreateWithFlags(&ev0, hipEventDisableTiming);
    if (e != hipSuccess) {
        group_free(b->grp);
        delete b;
        return fail(FDTD_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
    b->fork.push_back(ev0);
    // members' own streams may still hold work (uploads, resets)
    for (int i = 0; i < n; ++i) (void)hipStreamSynchronize(sims[i]->self.stream);
    int rc = batch_build(b);
    if (rc) {
        std::string keep = g_err;
        fdtd_batch_destroy(b);
        g_err = keep;
        return rc;
    }
    *out = b;
    return FDTD_OK;
}

extern "C" void fdtd_batch_destroy(FdtdBatch *b) {
    if (!b) return;
    (void)hipSetDevice(b->grp.device);
    for (hipStream_t st : b->streams) (void)hipStreamSynchronize(st);
    batch_clear_graphs(b);
    for (size_t l = 1; l < b->streams.size(); ++l) (void)hipStreamDestroy(b->streams[l]);
    for (hipEvent_t ev : b->fork) (void)hipEventDestroy(ev);
    group_free(b->grp);
    delete b;
}

extern "C" int fdtd_batch_set_launch(FdtdBatch *b, int ty, int zchunk, int xvec, int xwaves, int xcd_order) {
    CHECK(b, "NULL batch");
    CHECK(zchunk >= 0 && zchunk < 0x8000, "zchunk out of range");
    const Launch keep = b->grp.launch;
    b->grp.launch.ty = ty;
    b->grp.launch.zchunk = zchunk;
    b->grp.launch.xvec = xvec;
    b->grp.launch.xw = xwaves;
    b->grp.launch.xcd = xcd_order;
    int rc = batch_build(b);
    if (rc) {
        std::string err = g_err;
        b->grp.launch = keep;
        (void)batch_build(b);
        g_err = err;
    }
    return rc;
}

extern "C" int fdtd_batch_set_mode(FdtdBatch *b, int mode, int64_t small_cells, int lanes) {
    CHECK(b, "NULL batch");
    CHECK(mode == FDTD_BATCH_PACKED || mode == FDTD_BATCH_SERIAL || mode == FDTD_BATCH_HYBRID,
          "unknown batch mode %d", mode);
    CHECK(small_cells >= 0, "negative cell count");
    CHECK(lanes >= 1 && lanes <= FDTD_BATCH_MAX_LANES, "lanes %d outside [1, %d]", lanes, FDTD_BATCH_MAX_LANES);
    HIP_TRY(hipSetDevice(b->grp.device));
    for (hipStream_t st : b->streams) HIP_TRY(hipStreamSynchronize(st));
    const int keep_mode = b->mode, keep_lanes = b->nlanes;
    const long long keep_cells = b->small_cells;
    b->mode = mode;
    b->small_cells = small_cells;
    b->nlanes = lanes;
    int rc = batch_build(b);
    if (rc) {
        std::string err = g_err;
        b->mode = keep_mode;
        b->small_cells = keep_cells;
        b->nlanes = keep_lanes;
        (void)batch_build(b);
        g_err = err;
    }
    return rc;
}

extern "C" int fdtd_batch_run(FdtdBatch *b, int nsteps, int use_graph) {
    CHECK(b, "NULL batch");
    return batch_run(b, nsteps, use_graph != 0);
}

extern "C" int fdtd_batch_busy(FdtdBatch *b) {
    CHECK(b, "NULL batch");
    HIP_TRY(hipSetDevice(b->grp.device));
    const hipError_t e = hipStreamQuery(b->grp.stream);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return fail(FDTD_EHIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

extern "C" int fdtd_batch_sync(FdtdBatch *b) {
    CHECK(b, "NULL batch");
    HIP_TRY(hipSetDevice(b->grp.device));
    HIP_TRY(hipStreamSynchronize(b->grp.stream));
    return FDTD_OK;
}

extern "C" int fdtd_batch_time_steps(FdtdBatch *b, int nsteps, int use_graph, float *ms_total) {
    CHECK(b, "NULL batch");
    CHECK(nsteps > 0, "nsteps must be positive");
    if (b->mode == FDTD_BATCH_PACKED)
        return time_group(b->grp, nsteps, use_graph != 0, ms_total, nullptr, nullptr, nullptr);
    HIP_TRY(hipSetDevice(b->grp.device));
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    HIP_TRY(hipEventRecord(ev[0], b->grp.stream));
    int rc = batch_run(b, nsteps, use_graph != 0);
    if (rc == FDTD_OK) {
        HIP_TRY(hipEventRecord(ev[1], b->grp.stream));
        HIP_TRY(hipEventSynchronize(ev[1]));
        float tot = 0;
        HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev[1]));
        if (ms_total) *ms_total = tot;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}
