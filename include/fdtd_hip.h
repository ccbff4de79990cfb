/*
 * fdtd_hip.h — C-ABI of libfdtd_hip.so, the MI355X (gfx950) FDTD engine.
 *
 * The reference has no C-ABI on this path: its backends are Python classes
 * with the protocol
 *     __init__(params, mpi=False)                 fdtdsarsim/gpu/main_sim.py:45
 *     run(Ca, Cb, material_name, material_meta)   fdtdsarsim/gpu/main_sim.py:74
 * driven by pycuda launches of 23 runtime-compiled kernels
 * (gpu/main_sim.py:216-355, gpu/field_updates.py:9-907).  These entry points
 * are what such a backend needs from native code, and what a ctypes binding
 * (sar_pixel_simulator_amd/_native.py) or any other FFI binds.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative FDTD_E* code on
 *     failure; fdtd_last_error() then returns a thread-local message.
 *   - Host arrays are C-order float32 (nz, ny, nx), x fastest, as the
 *     reference's numpy arrays; the library copies them and never keeps the
 *     caller's pointer.  Complex outputs are interleaved (re, im) float32
 *     (numpy complex64) or float64 (complex128).
 *   - Handles belong to the library until destroyed/released.  One handle is
 *     used by one host thread at a time; different handles are independent.
 *   - Work is queued on the library's HIP stream for the handle's device;
 *     fdtd_sim_run() is asynchronous, fdtd_sim_sync() waits.
 */
#ifndef FDTD_HIP_H
#define FDTD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_ABI_VERSION 3
#define FDTD_MAX_PML 16

/* error codes */
#define FDTD_OK 0
#define FDTD_EINVAL -1    /* bad argument, shape or state */
#define FDTD_EHIP -2      /* HIP runtime error */
#define FDTD_ENOMEM -3    /* device allocation failed / over budget */
#define FDTD_ESTATE -4    /* call not valid in the handle's state */

/* simulation kinds */
#define FDTD_MODE_MAIN 0      /* H, E, phasors; TFSF injection (cpu/main_sim.py:140-195) */
#define FDTD_MODE_INCIDENT 1  /* E, hard source, H, record, incident phasors (cpu/incident_sim.py:134-161) */

/* phasor accumulation */
#define FDTD_DFT_COMPENSATED 0  /* float-float running sum per sample */
#define FDTD_DFT_REFERENCE 1    /* complex64 running sum rounded every step, as cpu/field_updates.py:385-400 */

/* Geometry and run length of one simulation.  Field names follow GeomParams
 * (fdtdsarsim/parameters.py:57-86). */
typedef struct FdtdGeom {
    int32_t nx, ny, nz;     /* size_x, size_y, size_z */
    int32_t npml;           /* size_pml (<= FDTD_MAX_PML) */
    int32_t tfsf_z;         /* MAIN: TFSF plane; INCIDENT: feed-in plane */
    int32_t read_out;       /* INCIDENT: read-out plane; MAIN: unused */
    int32_t nt;             /* steps the simulation runs (size_t / size_t_inc) */
    int32_t nt_inc;         /* MAIN: injected steps (size_t_inc); INCIDENT: source steps (size_t_pulse) */
    int32_t mode;           /* FDTD_MODE_* */
    int32_t dft_accum;      /* FDTD_DFT_* */
} FdtdGeom;

/* Update coefficients.  Db, dtm0, dx, dte0dx are the float32 scalars of
 * cpu/main_sim.py:119-122; omega, dt the float64 values of :101-102; the
 * eight CPML profiles are Absorbers' 1-D b/c profiles (absorbers.py:121-153)
 * in the order be_dlf, ce_dlf, bh_dlf, ch_dlf, be_urb, ce_urb, bh_urb, ch_urb,
 * each npml long (entries past npml ignored). */
typedef struct FdtdCoeffs {
    float Db, dtm0, dx, dte0dx;
    double omega, dt;
    double cell_size;       /* float64 dx [m] for the near-to-far transform (processresults.py:197) */
    float prof[8][FDTD_MAX_PML];
} FdtdCoeffs;

typedef struct FdtdDeviceInfo {
    int32_t device;
    int32_t cu_count;
    int32_t xcd_count;
    int32_t warp_size;
    int64_t total_mem;
    int64_t free_mem;
    char arch[64];
    char name[128];
} FdtdDeviceInfo;

typedef struct FdtdSim FdtdSim;
typedef struct FdtdBeam FdtdBeam;
typedef struct FdtdBatch FdtdBatch;

/* ---- library ---------------------------------------------------------- */
int fdtd_abi_version(void);
const char *fdtd_last_error(void);
/* Select the device for subsequent handle creation on this thread
 * (replaces experiment.py:33's CUDA_DEVICE and pycuda.autoinit). */
int fdtd_init(int device);
int fdtd_device_info(FdtdDeviceInfo *out);

/* ---- one simulation (MainSimGpu's device state, gpu/main_sim.py:512-787) */
/* Bytes of device memory a simulation of this geometry holds. */
int fdtd_sim_footprint(const FdtdGeom *g, int64_t *bytes);
int fdtd_sim_create(const FdtdGeom *g, const FdtdCoeffs *c, FdtdSim **out);
void fdtd_sim_destroy(FdtdSim *s);
/* Per-cell Ca, Cb (nz, ny, nx) float32 (gpu/main_sim.py:203-204). */
int fdtd_sim_set_materials(FdtdSim *s, const float *Ca, const float *Cb);
/* Same value in every cell (the incident simulation's vacuum,
 * cpu/incident_sim.py:115-116). */
int fdtd_sim_set_uniform_material(FdtdSim *s, float ca, float cb);
/* ---- material models built on the device ------------------------------ */
/* No counterpart in the reference, whose generators fill (eps_r, sigma)
 * volumes on the host (materials/soilmodels.py:16-298,
 * materials/materials.py:112-160) and ship Ca, Cb to the device
 * (gpu/main_sim.py:203-204).  Here a model is described by its layers,
 * inclusions and int32 surface maps; the device fills eps_r and sigma per
 * cell and forms Ca, Cb with compute_material_arrays' float32 arithmetic
 * (materials/materials.py:23-68), so a request ships O(nx ny) bytes. */
#define FDTD_MAX_LAYERS 4
enum {
    FDTD_LAYER_CONST = 0,     /* constant (eps, sigma) */
    FDTD_LAYER_RANDOM = 1,    /* eps ~ N(eps, eps_std^2) clipped at 1, sigma ~ U(sigma, sigma_hi) per cell */
    FDTD_LAYER_GRADIENT = 2   /* Hallikainen soil, moisture linear in z between the layer below's
                                 surface and this layer's (soilmodels.py:233-247) */
};

typedef struct FdtdMatLayer {
    int32_t kind;
    uint32_t seed;            /* RANDOM: stream key of the per-cell draws */
    double eps, sigma;        /* CONST: eps_r, sigma [S/m]; RANDOM: eps mean, sigma lower bound */
    double eps_std, sigma_hi; /* RANDOM */
    double mv_bottom, mv_top; /* GRADIENT: moisture at the lower and at this surface */
    double freq;              /* GRADIENT: frequency [Hz] */
    double eps0;              /* GRADIENT: eps0 of sigma = |eps''| eps0 2 pi f (scipy's value) */
    double f0, f1;            /* GRADIENT: Hallikainen table frequencies around freq */
    double poly[2][2][3];     /* GRADIENT: [f0, f1][real, imaginary][a, b, c] of eps = a + b mv + c mv^2 */
} FdtdMatLayer;

typedef struct FdtdMatCylinder {
    double c[3], u[3];        /* centre [m] and unit axis, (z, y, x) */
    double radius, length, eps, sigma;
    int32_t lo[3], hi[3];     /* cell box [lo, hi) that is tested, (z, y, x) */
} FdtdMatCylinder;

typedef struct FdtdMatSpec {
    int32_t nz_mat;           /* material planes; vacuum above (materials.py:42-47) */
    int32_t nlayers;          /* 0..FDTD_MAX_LAYERS */
    int32_t ncyl;
    int32_t vox_z1;           /* Bernoulli(vox_frac) inclusions in planes z < vox_z1 */
    double vox_frac, vox_eps, vox_sigma;
    uint32_t vox_seed, pad0;
    float dt, dx, eps0, pad1; /* float32 time step, cell size and eps0 of compute_material_arrays */
    FdtdMatLayer layer[FDTD_MAX_LAYERS];
} FdtdMatSpec;

/* Build Ca, Cb of a MAIN simulation on the device.  Per cell, in order:
 * vacuum; a voxel inclusion where z < vox_z1 and its draw is below
 * vox_frac; the last cylinder whose axis segment lies within length/2 and
 * radius of the cell centre (z dx, y dx, x dx); the first layer k with
 * z <= surfaces[k][y][x].  Per-cell draws are SplitMix64 hashes of
 * (seed, cell index, draw), normals by Box-Muller.  surfaces: nlayers
 * (ny, nx) int32 maps; cyl: ncyl cylinders (host arrays, copied). */
int fdtd_sim_build_materials(FdtdSim *s, const FdtdMatSpec *spec, const int32_t *surfaces,
                             const FdtdMatCylinder *cyl);
/* No counterpart in the reference (its arrays stay on the host).  Copy the
 * simulation's Ca, Cb, each (nz, ny, nx) float32, to the host. */
int fdtd_sim_get_materials(FdtdSim *s, float *Ca, float *Cb);
/* INCIDENT: hard-source planes, ex and ey each (n_src, ny, nx) float32
 * (imaginary parts of the excitation, cpu/incident_sim.py:110-111,142-144). */
int fdtd_sim_set_source(FdtdSim *s, const float *src_ex, const float *src_ey, int n_src);
/* MAIN: beam injected at tfsf_z; INCIDENT: beam the read-out planes are
 * recorded into.  The sim keeps a reference until destroyed or re-attached. */
int fdtd_sim_attach_beam(FdtdSim *s, FdtdBeam *b);
/* Zero fields, CPML state and phasors; step counter to 0
 * (gpu/main_sim.py:691-787). */
int fdtd_sim_reset(FdtdSim *s);
/* Queue nsteps time steps (asynchronous).  Steps past geom.nt are no-ops. */
int fdtd_sim_run(FdtdSim *s, int nsteps);
int fdtd_sim_sync(FdtdSim *s);
int fdtd_sim_step_count(FdtdSim *s, int *t);
/* Phasor `index` in the reference's save order (cpu/main_sim.py:220-229:
 * ex_ph_d ey_ph_d hx_ph_d hy_ph_d, ..._u, ey_ph_l ez_ph_l hy_ph_l hz_ph_l,
 * ..._r, ex_ph_f ez_ph_f hx_ph_f hz_ph_f, ..._b) as complex64 with the
 * reference's shape.  INCIDENT: indices 0..3 are ex/ey/hx/hy_inc_phasor. */
int fdtd_sim_get_phasor(FdtdSim *s, int index, float *out_c64);
/* Field component 0..5 = ex ey ez hx hy hz, (nz, ny, nx) float32. */
int fdtd_sim_get_field(FdtdSim *s, int comp, float *out);
/* Overwrite a field component from host (nz, ny, nx) float32 (tests). */
int fdtd_sim_set_field(FdtdSim *s, int comp, const float *in);
/* CPML state psi 0..23 in the reference's shapes (cpu/fdtd_sim.py:165-187):
 * E d,u: Exz Eyz; l,r: Eyx Ezx; f,b: Exy Ezy; then H likewise. */
int fdtd_sim_get_psi(FdtdSim *s, int index, float *out);
/* No counterpart in the reference (its step is 26 fixed launches,
 * gpu/main_sim.py:216-355).  MAIN only, before the first step: 0 = split
 * stepper (an H kernel, then an E kernel, in place; the library default);
 * 1 = fused stepper on row strips (one kernel computes H then E per step,
 * one wavefront walks a strip's rows; fields double-buffered); 2 = fused
 * stepper with one row per wavefront (ty = 3 or 7 rows per block, xvec 1 or
 * 2); 3 = halo stepper: one kernel per step on the split stepper's tiles
 * (one row per wavefront, up to 8 wavefronts per block, xvec 1 or 2), each
 * wavefront recomputing the H of the row below, of its chunk's first plane
 * and of the cell left of it instead of reading it (fastest from about 64M
 * cells).  All give identical results. */
int fdtd_sim_set_stepper(FdtdSim *s, int fused);
/* No counterpart in the reference.  Tile shape: each lane owns xvec
 * consecutive x cells (1, 2, 4) and
 * marches z over zchunk cells (0 = whole column).  Split stepper: blocks of
 * xwaves x ty wavefronts (at most 4), xwaves (1, 2, 4) side by side along
 * x on each of ty (1, 2, 4) rows; halo stepper likewise, at most 8
 * wavefronts (ty also 8).  Fused stepper: xwaves 1, ty 2, 4 or 8
 * rows per wavefront with (4 ty + 2) xvec <= 64 (LDS).  0 for ty, xvec or
 * xwaves: chosen from the grid (flat tiles, below, where they fill the lanes
 * better).  xwaves -1 (split stepper, xvec 0 or 1, ty 0, plain order or
 * fine bands, row pitch >= 64): flat tiles, blocks of 4 wavefronts each over
 * 64 consecutive cells of a plane's row-major (y, x) index, so that narrow
 * rows share wavefronts instead of each being padded to whole ones.  xcd_order:
 * 0 runs tiles in plain order, 1
 * deals each XCD a contiguous range of tiles, 2 (split and halo only) deals
 * each XCD a band of rows within every z chunk, 3 + b (split and halo only,
 * b = 0..3) deals the XCDs bands of 2^b row blocks in turn, so that an
 * XCD's consecutive blocks are x and y neighbours.  An invalid shape is
 * rejected and the previous one kept. */
int fdtd_sim_set_launch(FdtdSim *s, int ty, int zchunk, int xvec, int xwaves, int xcd_order);
/* No counterpart in the reference.  Run nsteps and report HIP-event times
 * on the library stream: whole run and per-kernel averages (ms), the latter
 * from whole-domain launches; with the fused stepper ms_h is the fused
 * kernel and ms_e is 0.  Any pointer may be NULL. */
int fdtd_sim_time_steps(FdtdSim *s, int nsteps, float *ms_total, float *ms_h,
                        float *ms_e, float *ms_dft);

/* Near-to-far transform of the top-face phasors (near2far.py:28-156), FP64
 * on the device.  thetas in degrees; out = E_theta, E_phi, H_theta, H_phi,
 * each ntheta complex128 (interleaved), i.e. 8*ntheta doubles. */
int fdtd_sim_near2far(FdtdSim *s, const double *thetas_deg, int ntheta,
                      double phi_deg, double r, double freq, double *out);
/* No counterpart in the reference.  fdtd_sim_near2far without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): its work
 * area (angle list, phase tables, partial sums, output staging) is a
 * preallocated per-device array of FDTD_N2F_SCRATCH_BYTES, held by each GPU
 * the library runs on until the process ends.  Same arguments and results
 * (bit for bit) as fdtd_sim_near2far; rejects ntheta > FDTD_N2F_MAX_THETA or
 * a phasor plane wider than FDTD_N2F_MAX_SIDE samples with FDTD_EINVAL.
 * Every call on one GPU shares that GPU's work area: calls that overlap on
 * one GPU (several host threads) race and must be serialised by the caller. */
#define FDTD_N2F_MAX_THETA 181
#define FDTD_N2F_MAX_SIDE 2047
#define FDTD_N2F_SCRATCH_BYTES 12621568
int fdtd_sim_near2far_noalloc(FdtdSim *s, const double *thetas_deg, int ntheta,
                              double phi_deg, double r, double freq, double *out);
/* Same transform of caller-supplied complex64 top-face phasors, each
 * (ny_ph, nx_ph): ex_u, ey_u, hx_u, hy_u. */
int fdtd_near2far_host(const float *ex_u, const float *ey_u, const float *hx_u,
                       const float *hy_u, int nx_ph, int ny_ph, int nz_ph,
                       double dx, const double *thetas_deg, int ntheta,
                       double phi_deg, double r, double freq, double *out);
/* The transform over a theta x phi grid in one launch, for a phi array as
 * FarFieldTransform.__call__ takes it (near2far.py:28-112, its phi loop):
 * out = E_theta, E_phi, H_theta, H_phi, each (nphi, ntheta) complex128,
 * i.e. 8*nphi*ntheta doubles. */
int fdtd_sim_near2far_grid(FdtdSim *s, const double *thetas_deg, int ntheta,
                           const double *phis_deg, int nphi, double r,
                           double freq, double *out);
int fdtd_near2far_host_grid(const float *ex_u, const float *ey_u,
                            const float *hx_u, const float *hy_u, int nx_ph,
                            int ny_ph, int nz_ph, double dx,
                            const double *thetas_deg, int ntheta,
                            const double *phis_deg, int nphi, double r,
                            double freq, double *out);

/* ---- incident beam (the (nt_inc, ny, nx) ex/ey/hx/hy datasets of the
 * incident HDF5 file, cpu/incident_sim.py:95-98), resident in HBM --------- */
/* ex, ey, hx, hy each (nt, ny, nx) float32; all NULL = zero-filled beam to
 * record into. */
int fdtd_beam_create(int nt, int ny, int nx, const float *ex, const float *ey,
                     const float *hx, const float *hy, FdtdBeam **out);
/* Overwrite steps [t0, t0 + n) of every component from host planes, each
 * (n, ny, nx) float32. */
int fdtd_beam_upload(FdtdBeam *b, int t0, int n, const float *ex, const float *ey,
                     const float *hx, const float *hy);
int fdtd_beam_get(FdtdBeam *b, int comp, float *out);
void fdtd_beam_release(FdtdBeam *b);

/* Tapered-wave beam description for fdtd_beam_synthesize. */
typedef struct FdtdBeamSpec {
    int32_t polarization;     /* 0 HH, 1 VV, 2 RHCP, 3 LHCP */
    float cb_vacuum;          /* Cb of the incident simulation's vacuum */
    double theta_deg, phi_deg;
    double gauss_radius;      /* [m]: illumination width / (2 sqrt(2 ln 10)) */
    const double *pulse_fft;  /* nt complex128 (re, im): FFT of the temporal pulse */
    const double *all_k;      /* nt wavenumbers 2 pi f_k / c of the FFT bins */
} FdtdBeamSpec;

/* Whole incident-beam pipeline on the device (IncidentBeam.setup's
 * generate branch, incidentbeam.py:51-60): excitation synthesis
 * (create_inc_excitation / taper_k, incidentbeam.py:206-433: spectra per
 * temporal bin, FP64 2-D FFTs, complex64 inverse FFT over time, peak
 * normalisation to 10), then the INCIDENT-mode simulation `g` driven by the
 * excitation's imaginary parts, recording geom.nt steps into a new beam.
 * c->cell_size is the spatial sample spacing.  inc_phasors_c64 receives the
 * four incident phasors, each (ny - 2 npml, nx - 2 npml) complex64. */
int fdtd_beam_synthesize(const FdtdGeom *g, const FdtdCoeffs *c, const FdtdBeamSpec *spec,
                         FdtdBeam **out, float *inc_phasors_c64);

/* ---- batches: many simulations stepped by one launch per kernel ------- */
/* No counterpart in the reference, which runs one model per process
 * (experiment.py:69-97, experiment_mpi.py:86-255).  The batch references
 * the sims (they must outlive it).  Sims may differ in size; they share a
 * mode and a stepper; each keeps its own step counter. */
int fdtd_batch_create(FdtdSim *const *sims, int n, FdtdBatch **out);
void fdtd_batch_destroy(FdtdBatch *b);
/* As fdtd_sim_set_launch, for every member.  With ty, xvec and xwaves all 0
 * (automatic, the default) the split stepper gives each member its own
 * block width. */
int fdtd_batch_set_launch(FdtdBatch *b, int ty, int zchunk, int xvec, int xwaves, int xcd_order);
/* Queue nsteps steps of every member; use_graph != 0 captures the step
 * sequence into a HIP graph on first use and replays it. */
int fdtd_batch_run(FdtdBatch *b, int nsteps, int use_graph);
int fdtd_batch_sync(FdtdBatch *b);
/* No counterpart in the reference.  1 while steps queued on the batch are
 * still running, 0 once they are done (the batch's stream is idle); does not
 * wait.  A serving loop uses it to do host work only while the device is
 * busy. */
int fdtd_batch_busy(FdtdBatch *b);
int fdtd_batch_time_steps(FdtdBatch *b, int nsteps, int use_graph, float *ms_total);
/* No counterpart in the reference.  How a batch steps its members:
 * FDTD_BATCH_PACKED (the default) puts every member in one tile list, one
 * launch per kernel per step for all of them; FDTD_BATCH_SERIAL runs each
 * member's own launches (its own tile shape), a chunk of steps per member,
 * one member after another on each of `lanes` concurrent streams (members
 * dealt largest first to the lane with the fewest cells) -- with
 * small_cells > 0 only the members below small_cells are dealt to lanes,
 * and the others follow one after another once those are done;
 * FDTD_BATCH_HYBRID packs the members with fewer than small_cells cells and
 * runs the others as in FDTD_BATCH_SERIAL with small_cells 0.  Launches fork from and join back into the
 * batch's stream (one graph per chunk when captured); fdtd_batch_sync waits
 * for all of them.  Results are identical in every mode. */
#define FDTD_BATCH_PACKED 0
#define FDTD_BATCH_SERIAL 1
#define FDTD_BATCH_HYBRID 2
#define FDTD_BATCH_MAX_LANES 8
int fdtd_batch_set_mode(FdtdBatch *b, int mode, int64_t small_cells, int lanes);

/* ---- measurement support (no counterpart in the reference) ------------ */
/* Copy nbytes between two fresh device buffers reps times (after one
 * untimed copy) with a plain streaming kernel, vec floats per thread (1 or
 * 4), and report the average HIP-event time of one copy.  The vec = 1 form
 * makes the stencil kernels' coalesced 4-byte accesses: its bandwidth is the
 * practical ceiling they are measured against, and its known byte count
 * calibrates the PMC byte counters. */
int fdtd_probe_copy(int64_t nbytes, int vec, int reps, float *ms_per_copy);
/* nread input and nwrite output arrays of n floats, each thread reading one
 * element of every input and writing one of every output (4 bytes per lane
 * per stream): the HBM access mix of a stencil kernel without its neighbour
 * reuse or arithmetic -- 6+3 (H half), 8+3 (E half), 8+6 (fused step), and
 * 1+1.  Reports the average HIP-event time of one pass. */
int fdtd_probe_streams(int64_t n, int nread, int nwrite, int reps, float *ms_per_pass);
/* No counterpart in the reference.  The stencil kernels' access pattern
 * without their arithmetic over an nx x ny x nz grid laid out like a
 * simulation's fields: blocks of 256 lanes along x, tiles in plain order, a
 * z march over zchunk planes carrying the plane-to-plane values in
 * registers, and the y +- 1 row loads.  kind 0: the split H half (9
 * algorithmic streams per cell), 1: the split E half (11), 2: the halo
 * stepper's fused step (14; the row below's E and H re-read, both buffer
 * sets).  Reports the average HIP-event time of one pass: the practical
 * ceiling of that pattern. */
int fdtd_probe_stencil(int nx, int ny, int nz, int zchunk, int kind, int reps, float *ms_per_pass);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_HIP_H */
