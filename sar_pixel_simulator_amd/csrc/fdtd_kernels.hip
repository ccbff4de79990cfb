// Stencil, phasor and near-to-far kernels of the MI355X FDTD engine.
//
// Must be compiled with -ffp-contract=off: the field and CPML updates keep
// the reference's one-rounding-per-operation arithmetic
// (fdtdsarsim/cpu/field_updates.py, numba without fastmath) so that results
// are bit-identical to the CPU path on the same inputs.
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

__device__ __forceinline__ int load_step(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Per-cell updates shared by the split and the fused steppers.
// ---------------------------------------------------------------------------

// E values around an H node (x, y, z) at the old time level.
struct EAround {
    float ex_c, ex_n, ex_yp;   // ex(z), ex(z+1), ex(y+1)
    float ey_c, ey_n, ey_xp;   // ey(z), ey(z+1), ey(x+1)
    float ez_c, ez_yp, ez_xp;  // ez(z), ez(y+1), ez(x+1)
};

// H half-step of one cell: Yee update (cpu/field_updates.py:198-240), the
// six H-CPML corrections in the reference order d,u,l,r,f,b (:243-366,
// cpu/fdtd_sim.py:104-133), then the E_inc injection at tfsf_z in MAIN mode
// (cpu/main_sim.py:147-150).  Every correction reads E only, so applying
// them in sequence per cell gives the values of the reference's separate
// passes exactly.  psi is read from psi_rd and, when psi_wr is not null,
// the updated value is written there (psi_wr == psi_rd updates in place).
__device__ __forceinline__ void h_cell(const Desc *__restrict__ d, int x, int y, int z, int t,
                                       const EAround &e, float &hx, float &hy, float &hz,
                                       float *const *psi_rd, float *const *psi_wr) {
    const int nx = d->nx, ny = d->ny, nz = d->nz, np = d->npml;
    const long long sy = d->sy;
    const float Db = d->Db, dtm0 = d->dtm0, dx = d->dx;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    if (z < nz - 1) {
        if (x >= 1 && xin && yin) hx -= Db * ((e.ez_yp - e.ez_c) - (e.ey_n - e.ey_c));
        if (y >= 1 && yin && xin) hy -= Db * ((e.ex_n - e.ex_c) - (e.ez_xp - e.ez_c));
        if (z >= 1 && yin && xin) hz -= Db * ((e.ey_xp - e.ey_c) - (e.ex_yp - e.ex_c));
        if (yin && xin) {
            if (z < np - 1) {                                       // down
                const float bh = d->prof[BH_DLF][z], ch = d->prof[CH_DLF][z];
                const long long p = (long long)z * d->du_sz + (long long)y * sy + x;
                const float a = (bh * psi_rd[pHyz_d][p]) + ((ch * (e.ex_n - e.ex_c)) / dx);
                const float b = (bh * psi_rd[pHxz_d][p]) + ((ch * (e.ey_n - e.ey_c)) / dx);
                if (psi_wr) { psi_wr[pHyz_d][p] = a; psi_wr[pHxz_d][p] = b; }
                hy -= dtm0 * a;
                hx += dtm0 * b;
            }
            if (z >= nz - np) {                                     // up
                const int zl = z - (nz - np);
                const float bh = d->prof[BH_URB][zl], ch = d->prof[CH_URB][zl];
                const long long p = (long long)zl * d->du_sz + (long long)y * sy + x;
                const float a = (bh * psi_rd[pHyz_u][p]) + ((ch * (e.ex_n - e.ex_c)) / dx);
                const float b = (bh * psi_rd[pHxz_u][p]) + ((ch * (e.ey_n - e.ey_c)) / dx);
                if (psi_wr) { psi_wr[pHyz_u][p] = a; psi_wr[pHxz_u][p] = b; }
                hy -= dtm0 * a;
                hx += dtm0 * b;
            }
        }
        if (yin) {
            if (x < np - 1) {                                       // left
                const float bh = d->prof[BH_DLF][x], ch = d->prof[CH_DLF][x];
                const long long p = (long long)z * d->lr_sz + (long long)y * np + x;
                const float a = (bh * psi_rd[pHzx_l][p]) + ((ch * (e.ey_xp - e.ey_c)) / dx);
                const float b = (bh * psi_rd[pHyx_l][p]) + ((ch * (e.ez_xp - e.ez_c)) / dx);
                if (psi_wr) { psi_wr[pHzx_l][p] = a; psi_wr[pHyx_l][p] = b; }
                hz -= dtm0 * a;
                hy += dtm0 * b;
            }
            if (x >= nx - np && xin) {                              // right
                const int xl = x - (nx - np);
                const float bh = d->prof[BH_URB][xl], ch = d->prof[CH_URB][xl];
                const long long p = (long long)z * d->lr_sz + (long long)y * np + xl;
                const float a = (bh * psi_rd[pHzx_r][p]) + ((ch * (e.ey_xp - e.ey_c)) / dx);
                const float b = (bh * psi_rd[pHyx_r][p]) + ((ch * (e.ez_xp - e.ez_c)) / dx);
                if (psi_wr) { psi_wr[pHzx_r][p] = a; psi_wr[pHyx_r][p] = b; }
                hz -= dtm0 * a;
                hy += dtm0 * b;
            }
        }
        if (xin) {
            if (y < np - 1) {                                       // front
                const float bh = d->prof[BH_DLF][y], ch = d->prof[CH_DLF][y];
                const long long p = (long long)z * d->fb_sz + (long long)y * sy + x;
                const float a = (bh * psi_rd[pHxy_f][p]) + ((ch * (e.ez_yp - e.ez_c)) / dx);
                const float b = (bh * psi_rd[pHzy_f][p]) + ((ch * (e.ex_yp - e.ex_c)) / dx);
                if (psi_wr) { psi_wr[pHxy_f][p] = a; psi_wr[pHzy_f][p] = b; }
                hx -= dtm0 * a;
                hz += dtm0 * b;
            }
            if (y >= ny - np && yin) {                              // back
                const int yl = y - (ny - np);
                const float bh = d->prof[BH_URB][yl], ch = d->prof[CH_URB][yl];
                const long long p = (long long)z * d->fb_sz + (long long)yl * sy + x;
                const float a = (bh * psi_rd[pHxy_b][p]) + ((ch * (e.ez_yp - e.ez_c)) / dx);
                const float b = (bh * psi_rd[pHzy_b][p]) + ((ch * (e.ex_yp - e.ex_c)) / dx);
                if (psi_wr) { psi_wr[pHxy_b][p] = a; psi_wr[pHzy_b][p] = b; }
                hx -= dtm0 * a;
                hz += dtm0 * b;
            }
        }
    }
    if (d->mode == FDTD_MODE_MAIN && t < d->nt_inc && z == d->tfsf_z) {
        const long long q = ((long long)t * ny + y) * nx + x;
        hx += Db * d->beam[1 * d->beam_cs + q];   // ey_inc
        hy -= Db * d->beam[0 * d->beam_cs + q];   // ex_inc
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
// top plane (z < nz - 1).
__device__ __forceinline__ void e_cell(const Desc *__restrict__ d, int x, int y, int z, int t, float ca,
                                       float cb, const HAround &h, float &ex, float &ey, float &ez) {
    const int nx = d->nx, ny = d->ny, nz = d->nz, np = d->npml;
    const long long sy = d->sy;
    const float dx = d->dx;
    const bool xlo = x >= 1, ylo = y >= 1, zlo = z >= 1;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    float *const *psi = d->psi[0];
    if (z < nz - 1) {
        if (zlo && ylo && yin && xin) ex = (ex * ca) + (cb * ((h.hz_c - h.hz_ym) - (h.hy_c - h.hy_zm)));
        if (zlo && yin && xlo && xin) ey = (ey * ca) + (cb * ((h.hx_c - h.hx_zm) - (h.hz_c - h.hz_xm)));
        if (ylo && yin && xlo && xin) ez = (ez * ca) + (cb * ((h.hy_c - h.hy_xm) - (h.hx_c - h.hx_ym)));
        const float cbdx = cb * dx;
        if (yin && xin) {
            if (zlo && z < np) {                                    // down
                const float be = d->prof[BE_DLF][z], ce = d->prof[CE_DLF][z];
                const long long p = (long long)z * d->du_sz + (long long)y * sy + x;
                const float a = (be * psi[pEyz_d][p]) + ((ce * (h.hx_c - h.hx_zm)) / dx);
                const float b = (be * psi[pExz_d][p]) + ((ce * (h.hy_c - h.hy_zm)) / dx);
                psi[pEyz_d][p] = a; psi[pExz_d][p] = b;
                ey += cbdx * a;
                ex -= cbdx * b;
            }
            if (z >= nz - np) {                                     // up
                const int zl = z - (nz - np);
                const float be = d->prof[BE_URB][zl], ce = d->prof[CE_URB][zl];
                const long long p = (long long)zl * d->du_sz + (long long)y * sy + x;
                const float a = (be * psi[pEyz_u][p]) + ((ce * (h.hx_c - h.hx_zm)) / dx);
                const float b = (be * psi[pExz_u][p]) + ((ce * (h.hy_c - h.hy_zm)) / dx);
                psi[pEyz_u][p] = a; psi[pExz_u][p] = b;
                ey += cbdx * a;
                ex -= cbdx * b;
            }
        }
        if (yin) {
            if (xlo && x < np) {                                    // left
                const float be = d->prof[BE_DLF][x], ce = d->prof[CE_DLF][x];
                const long long p = (long long)z * d->lr_sz + (long long)y * np + x;
                const float a = (be * psi[pEzx_l][p]) + ((ce * (h.hy_c - h.hy_xm)) / dx);
                const float b = (be * psi[pEyx_l][p]) + ((ce * (h.hz_c - h.hz_xm)) / dx);
                psi[pEzx_l][p] = a; psi[pEyx_l][p] = b;
                ez += cbdx * a;
                ey -= cbdx * b;
            }
            if (x >= nx - np && xin) {                              // right
                const int xl = x - (nx - np);
                const float be = d->prof[BE_URB][xl], ce = d->prof[CE_URB][xl];
                const long long p = (long long)z * d->lr_sz + (long long)y * np + xl;
                const float a = (be * psi[pEzx_r][p]) + ((ce * (h.hy_c - h.hy_xm)) / dx);
                const float b = (be * psi[pEyx_r][p]) + ((ce * (h.hz_c - h.hz_xm)) / dx);
                psi[pEzx_r][p] = a; psi[pEyx_r][p] = b;
                ez += cbdx * a;
                ey -= cbdx * b;
            }
        }
        if (xin) {
            if (ylo && y < np) {                                    // front
                const float be = d->prof[BE_DLF][y], ce = d->prof[CE_DLF][y];
                const long long p = (long long)z * d->fb_sz + (long long)y * sy + x;
                const float a = (be * psi[pExy_f][p]) + ((ce * (h.hz_c - h.hz_ym)) / dx);
                const float b = (be * psi[pEzy_f][p]) + ((ce * (h.hx_c - h.hx_ym)) / dx);
                psi[pExy_f][p] = a; psi[pEzy_f][p] = b;
                ex += cbdx * a;
                ez -= cbdx * b;
            }
            if (y >= ny - np && yin) {                              // back
                const int yl = y - (ny - np);
                const float be = d->prof[BE_URB][yl], ce = d->prof[CE_URB][yl];
                const long long p = (long long)z * d->fb_sz + (long long)yl * sy + x;
                const float a = (be * psi[pExy_b][p]) + ((ce * (h.hz_c - h.hz_ym)) / dx);
                const float b = (be * psi[pEzy_b][p]) + ((ce * (h.hx_c - h.hx_ym)) / dx);
                psi[pExy_b][p] = a; psi[pEzy_b][p] = b;
                ex += cbdx * a;
                ez -= cbdx * b;
            }
        }
    }
    if (z == d->tfsf_z) {
        if (d->mode == FDTD_MODE_MAIN) {
            if (t < d->nt_inc) {
                const long long q = ((long long)t * ny + y) * nx + x;
                ey += d->dte0dx * d->beam[2 * d->beam_cs + q];   // hx_inc
                ex -= d->dte0dx * d->beam[3 * d->beam_cs + q];   // hy_inc
            }
        } else if (t < d->n_src && ylo && yin && xin) {
            const long long q = ((long long)t * ny + y) * nx + x;
            ex = d->src[q];
            ey = d->src[d->src_cs + q];
        }
    }
}

// ---------------------------------------------------------------------------
// Split stepper: H half-step kernel, then E half-step kernel, fields updated
// in place (set 0).  Used for INCIDENT runs (E before H) and available for
// MAIN runs.  A block of 64 x ty lanes covers a (64 V) x ty tile: each lane
// owns V consecutive x cells of one row, loads and stores them as one
// V-wide access (16 bytes for V = 4), and marches z upward, reusing the
// plane above (H half: ex, ey) or below (E half: hx, hy) from registers.
// The x +- 1 neighbours come from the lane's own cells or, at its ends, from
// the adjacent lane by a wavefront shuffle (from memory at the tile edge).
// Rows are padded to a multiple of 16 floats, so a lane's V cells are
// either all inside the row or all in its padding.
// ---------------------------------------------------------------------------
template <int V>
__device__ __forceinline__ void vld(float (&r)[V], const float *p) {
    if constexpr (V == 4) {
        const float4 q = *reinterpret_cast<const float4 *>(p);
        r[0] = q.x; r[1] = q.y; r[2] = q.z; r[3] = q.w;
    } else if constexpr (V == 2) {
        const float2 q = *reinterpret_cast<const float2 *>(p);
        r[0] = q.x; r[1] = q.y;
    } else {
        r[0] = *p;
    }
}

template <int V>
__device__ __forceinline__ void vst(float *p, const float (&r)[V]) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(r[0], r[1], r[2], r[3]);
    } else if constexpr (V == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(r[0], r[1]);
    } else {
        *p = r[0];
    }
}

template <int V>
__device__ __forceinline__ void vset(float (&r)[V], float v) {
#pragma unroll
    for (int k = 0; k < V; ++k) r[k] = v;
}

template <int V>
__global__ void __launch_bounds__(256)
k_h_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles) {
    const int4 tl = tiles[xcd_tile(blockIdx.x, ntiles)];
    const Desc *__restrict__ d = descs + tl.x;
    const int t = load_step(d->step);
    if (t >= d->nt) return;                            // uniform per block
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int y = tl.z + threadIdx.y;
    if (y >= ny) return;                               // uniform per wavefront (one row)
    const int xb = tl.y + V * threadIdx.x;             // the lane's first cell
    const bool act = xb < nx;
    // the x + 1 neighbour of the lane's last cell is needed when that cell
    // is below nx - 1; for lane 63 it lies in the next tile
    const bool next_mem = threadIdx.x == kWave - 1 && xb + V < nx;
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const long long sy = d->sy, sz = d->sz;
    const float *__restrict__ Ex = d->f[0][0];
    const float *__restrict__ Ey = d->f[0][1];
    const float *__restrict__ Ez = d->f[0][2];
    float *__restrict__ Hx = d->f[0][3];
    float *__restrict__ Hy = d->f[0][4];
    float *__restrict__ Hz = d->f[0][5];
    float *const *psi = d->psi[0];
    const bool yin = y < ny - 1;

    long long o = (long long)z0 * sz + (long long)y * sy + xb;
    float exc[V], eyc[V];
    vset(exc, 0.f);
    vset(eyc, 0.f);
    if (act) {
        vld(exc, Ex + o);
        vld(eyc, Ey + o);
    }
    for (int z = z0; z < z1; ++z, o += sz) {
        const bool zin = z < nz - 1;
        float hx[V], hy[V], hz[V], exn[V], eyn[V], ezc[V], ezyp[V], exyp[V];
        vset(exn, 0.f);
        vset(eyn, 0.f);
        vset(ezc, 0.f);
        vset(ezyp, 0.f);
        vset(exyp, 0.f);
        if (act) {
            vld(hx, Hx + o);
            vld(hy, Hy + o);
            vld(hz, Hz + o);
            if (zin) {
                vld(exn, Ex + o + sz);
                vld(eyn, Ey + o + sz);
                vld(ezc, Ez + o);
                if (yin) {
                    vld(ezyp, Ez + o + sy);
                    vld(exyp, Ex + o + sy);
                }
            }
        }
        float ey_next = __shfl_down(eyc[0], 1, kWave);
        float ez_next = __shfl_down(ezc[0], 1, kWave);
        if (next_mem && zin) {
            ey_next = Ey[o + V];
            ez_next = Ez[o + V];
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (xb + k < nx) {
                EAround e;
                e.ex_c = exc[k]; e.ex_n = exn[k]; e.ex_yp = exyp[k];
                e.ey_c = eyc[k]; e.ey_n = eyn[k];
                e.ey_xp = k + 1 < V ? eyc[k + 1 < V ? k + 1 : 0] : ey_next;
                e.ez_c = ezc[k]; e.ez_yp = ezyp[k];
                e.ez_xp = k + 1 < V ? ezc[k + 1 < V ? k + 1 : 0] : ez_next;
                h_cell(d, xb + k, y, z, t, e, hx[k], hy[k], hz[k], psi, psi);
            }
        }
        if (act) {
            vst(Hx + o, hx);
            vst(Hy + o, hy);
            vst(Hz + o, hz);
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            exc[k] = exn[k];
            eyc[k] = eyn[k];
        }
    }
}

template <int V>
__global__ void __launch_bounds__(256)
k_e_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles) {
    const int4 tl = tiles[xcd_tile(blockIdx.x, ntiles)];
    const Desc *__restrict__ d = descs + tl.x;
    const int t = load_step(d->step);
    if (t >= d->nt) return;                            // uniform per block
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int y = tl.z + threadIdx.y;
    if (y >= ny) return;                               // uniform per wavefront (one row)
    const int xb = tl.y + V * threadIdx.x;
    const bool act = xb < nx;
    const bool first_lane = threadIdx.x == 0;
    const bool prev_mem = first_lane && xb >= 1 && act;   // x - 1 lies in the previous tile
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const long long sy = d->sy, sz = d->sz;
    float *__restrict__ Ex = d->f[0][0];
    float *__restrict__ Ey = d->f[0][1];
    float *__restrict__ Ez = d->f[0][2];
    const float *__restrict__ Hx = d->f[0][3];
    const float *__restrict__ Hy = d->f[0][4];
    const float *__restrict__ Hz = d->f[0][5];
    const bool ylo = y >= 1;
    const int uni_z = d->uni_z;
    const float ca_u = d->ca_u, cb_u = d->cb_u;

    long long o = (long long)z0 * sz + (long long)y * sy + xb;
    float hxzm[V], hyzm[V];
    vset(hxzm, 0.f);
    vset(hyzm, 0.f);
    if (act && z0 >= 1) {
        vld(hxzm, Hx + o - sz);
        vld(hyzm, Hy + o - sz);
    }
    for (int z = z0; z < z1; ++z, o += sz) {
        const bool zin = z < nz - 1;
        const bool mat = zin && z < uni_z;             // plane holds per-cell Ca/Cb
        float hxc[V], hyc[V], hzc[V], hzym[V], hxym[V], ex[V], ey[V], ez[V], ca[V], cb[V];
        vset(hxc, 0.f);
        vset(hyc, 0.f);
        vset(hzc, 0.f);
        vset(hzym, 0.f);
        vset(hxym, 0.f);
        vset(ca, ca_u);
        vset(cb, cb_u);
        if (act) {
            vld(hxc, Hx + o);
            vld(hyc, Hy + o);
            vld(ex, Ex + o);
            vld(ey, Ey + o);
            vld(ez, Ez + o);
            if (zin) {
                vld(hzc, Hz + o);
                if (ylo) {
                    vld(hzym, Hz + o - sy);
                    vld(hxym, Hx + o - sy);
                }
                if (mat) {
                    vld(ca, d->ca + o);
                    vld(cb, d->cb + o);
                }
            }
        }
        float hz_prev = __shfl_up(hzc[V - 1], 1, kWave);
        float hy_prev = __shfl_up(hyc[V - 1], 1, kWave);
        if (first_lane) {
            hz_prev = 0.f;
            hy_prev = 0.f;
            if (prev_mem && zin) {
                hz_prev = Hz[o - 1];
                hy_prev = Hy[o - 1];
            }
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (xb + k < nx) {
                HAround h;
                h.hx_c = hxc[k]; h.hy_c = hyc[k]; h.hz_c = hzc[k];
                h.hx_zm = hxzm[k]; h.hy_zm = hyzm[k];
                h.hz_ym = hzym[k]; h.hx_ym = hxym[k];
                h.hz_xm = k > 0 ? hzc[k > 0 ? k - 1 : 0] : hz_prev;
                h.hy_xm = k > 0 ? hyc[k > 0 ? k - 1 : 0] : hy_prev;
                e_cell(d, xb + k, y, z, t, ca[k], cb[k], h, ex[k], ey[k], ez[k]);
            }
        }
        if (act) {
            vst(Ex + o, ex);
            vst(Ey + o, ey);
            vst(Ez + o, ez);
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            hxzm[k] = hxc[k];
            hyzm[k] = hyc[k];
        }
    }
}

template __global__ void k_h_half<1>(const Desc *, const int4 *, int);
template __global__ void k_h_half<2>(const Desc *, const int4 *, int);
template __global__ void k_h_half<4>(const Desc *, const int4 *, int);
template __global__ void k_e_half<1>(const Desc *, const int4 *, int);
template __global__ void k_e_half<2>(const Desc *, const int4 *, int);
template __global__ void k_e_half<4>(const Desc *, const int4 *, int);

// ---------------------------------------------------------------------------
// Fused stepper (MAIN): one kernel per time step computes H(t+1/2) and then
// E(t+1) plane by plane, so each field value is read once and written once
// per step (56 B/cell instead of the split stepper's 80).  Fields and H-CPML
// state are double-buffered: step t reads set t&1 and writes set 1-(t&1).
//
// A block of 64 x TY threads owns a 63 x (TY-1) tile: lane 0 and row 0 are
// a one-cell halo whose H values (needed by E at x-1, y-1) are recomputed
// from the old level rather than exchanged between blocks; the owner tile
// computes the same values and is the only one that writes them.  Each
// thread holds one (x, y) column and marches z upward: E(z+1) read at plane
// z is E(z) at plane z+1, H(z) computed at plane z is H(z-1) at plane z+1.
// H at x-1 comes from the neighbouring lane (wavefront shuffle); H at y-1
// from the row below through LDS, one barrier per plane.  A tile starting
// at z0 > 0 recomputes H at z0-1 (not stored) for its first E plane.
// ---------------------------------------------------------------------------
template <int TY>
__global__ void __launch_bounds__(kWave * TY)
k_step_fused(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles) {
    __shared__ float lds_hz[2][TY][kWave];
    __shared__ float lds_hx[2][TY][kWave];
    const int4 tl = tiles[xcd_tile(blockIdx.x, ntiles)];
    const Desc *__restrict__ d = descs + tl.x;
    const int t = load_step(d->step);
    if (t >= d->nt) return;                            // uniform per block
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int x = tl.y - 1 + lx, y = tl.z - 1 + ly;    // lane 0 / row 0: halo
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;
    const int p = t & 1;
    const long long sy = d->sy, sz = d->sz;
    const float *__restrict__ Exo = d->f[p][0];
    const float *__restrict__ Eyo = d->f[p][1];
    const float *__restrict__ Ezo = d->f[p][2];
    const float *__restrict__ Hxo = d->f[p][3];
    const float *__restrict__ Hyo = d->f[p][4];
    const float *__restrict__ Hzo = d->f[p][5];
    float *__restrict__ Exn = d->f[1 - p][0];
    float *__restrict__ Eyn = d->f[1 - p][1];
    float *__restrict__ Ezn = d->f[1 - p][2];
    float *__restrict__ Hxn = d->f[1 - p][3];
    float *__restrict__ Hyn = d->f[1 - p][4];
    float *__restrict__ Hzn = d->f[1 - p][5];
    const bool ok = x >= 0 && x < nx && y >= 0 && y < ny;
    const bool own = ok && lx >= 1 && ly >= 1;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    const bool last_lane = lx == kWave - 1;
    const int zs = z0 > 0 ? z0 - 1 : 0;

    long long o = (long long)zs * sz + (long long)y * sy + x;
    EAround e;
    e.ex_c = ok ? Exo[o] : 0.f;
    e.ey_c = ok ? Eyo[o] : 0.f;
    float hx_zm = 0.f, hy_zm = 0.f;
    for (int z = zs; z < z1; ++z, o += sz) {
        const int buf = z & 1;
        const bool zin = z < nz - 1;
        float hx = 0.f, hy = 0.f, hz = 0.f;
        e.ex_n = e.ey_n = e.ez_c = e.ez_yp = e.ex_yp = 0.f;
        if (ok) {
            hx = Hxo[o];
            hy = Hyo[o];
            hz = Hzo[o];
            e.ez_c = Ezo[o];
            if (zin) {
                e.ex_n = Exo[o + sz];
                e.ey_n = Eyo[o + sz];
                if (yin) {
                    e.ez_yp = Ezo[o + sy];
                    e.ex_yp = Exo[o + sy];
                }
            }
        }
        e.ey_xp = __shfl_down(e.ey_c, 1, kWave);
        e.ez_xp = __shfl_down(e.ez_c, 1, kWave);
        if (last_lane) {                               // x+1 belongs to the next tile
            const bool okx = ok && xin;
            e.ey_xp = okx ? Eyo[o + 1] : 0.f;
            e.ez_xp = okx ? Ezo[o + 1] : 0.f;
        }
        const bool store = z >= z0;
        if (ok) h_cell(d, x, y, z, t, e, hx, hy, hz, d->psi[p], (own && store) ? d->psi[1 - p] : nullptr);
        lds_hz[buf][ly][lx] = hz;
        lds_hx[buf][ly][lx] = hx;
        const float hz_xm = __shfl_up(hz, 1, kWave);
        const float hy_xm = __shfl_up(hy, 1, kWave);
        __syncthreads();
        if (own && store) {
            HAround h;
            h.hx_c = hx; h.hy_c = hy; h.hz_c = hz;
            h.hx_zm = hx_zm; h.hy_zm = hy_zm;
            h.hz_ym = lds_hz[buf][ly - 1][lx];
            h.hx_ym = lds_hx[buf][ly - 1][lx];
            h.hz_xm = hz_xm; h.hy_xm = hy_xm;
            float ex = e.ex_c, ey = e.ey_c, ez = e.ez_c;
            float ca = d->ca_u, cb = d->cb_u;
            if (zin && z < d->uni_z) {
                ca = d->ca[o];
                cb = d->cb[o];
            }
            e_cell(d, x, y, z, t, ca, cb, h, ex, ey, ez);
            Exn[o] = ex;
            Eyn[o] = ey;
            Ezn[o] = ez;
            Hxn[o] = hx;
            Hyn[o] = hy;
            Hzn[o] = hz;
        }
        hx_zm = hx;
        hy_zm = hy;
        e.ex_c = e.ex_n;
        e.ey_c = e.ey_n;
    }
}

template __global__ void k_step_fused<8>(const Desc *, const int4 *, int);
template __global__ void k_step_fused<16>(const Desc *, const int4 *, int);

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
// k_advance then moves every member's step counter to the next step.
// ---------------------------------------------------------------------------
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
    const int t = load_step(d->step);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
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
                float4 r = v;
                if (c < 2) acc_ref(r, s[c], 0.5, w.x, w.y);
                else acc_ref(r, s[c], 0.25, w.z, w.w);
                v = r;
            }
        } else {
            const float4 we = d->tw32[2 * t], wh = d->tw32[2 * t + 1];
            for (int c = 0; c < 4; ++c) {
                float4 &v = acc[d->acc_off[4 * f + c] + j];
                float4 r = v;
                acc_comp(r, s[c], c < 2 ? we : wh);
                v = r;
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

// Advance every member's step counter after the step's kernels (a separate
// launch, so no kernel of the step reads a counter another has changed).
__global__ void __launch_bounds__(64)
k_advance(const Desc *__restrict__ descs, int nmodels) {
    for (int m = threadIdx.x; m < nmodels; m += blockDim.x) {
        int *sp = descs[m].step;
        const int t = load_step(sp);
        if (t < descs[m].nt) __hip_atomic_store(sp, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// ---------------------------------------------------------------------------
// Near-to-far transform over the top face (near2far.py:28-156), FP64.
// The integrand is linear in J = (-hy_u, hx_u) and M = (ey_u, -ex_u), so
// each theta needs the four Simpson-weighted sums of J_x, J_y, M_x, M_y
// times exp(jk(x sin th cos ph + y sin th sin ph + z_top cos th)); the phase
// factorises into per-column and per-row tables built on the host.
// Block (theta, split) reduces its rows with wavefront shuffles, then LDS.
// ---------------------------------------------------------------------------
struct cd { double re, im; };
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

__global__ void __launch_bounds__(256)
k_n2f_partial(const float2 *__restrict__ ex_u, const float2 *__restrict__ ey_u,
              const float2 *__restrict__ hx_u, const float2 *__restrict__ hy_u,
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
            const float2 a = hy_u[q], b = hx_u[q], c = ey_u[q], e = ex_u[q];
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
