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
// H half-step: Yee H update (cpu/field_updates.py:198-240), the six H-CPML
// corrections in the reference order d,u,l,r,f,b (:243-366,
// cpu/fdtd_sim.py:104-133) and, in MAIN mode, the E_inc injection at tfsf_z
// (cpu/main_sim.py:147-150) -- fused per cell.  Every correction of a cell
// reads E only, so applying them in sequence per cell gives the values of
// the reference's separate passes exactly.
//
// One thread owns an (x, y) column of a tile and marches z upward; ex(z+1)
// and ey(z+1) loaded at plane z are reused from registers at plane z+1.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_h_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles) {
    const int4 tl = tiles[xcd_tile(blockIdx.x, ntiles)];
    const Desc *__restrict__ d = descs + tl.x;
    const int t = load_step(d->step);
    if (t >= d->nt) return;
    const int nx = d->nx, ny = d->ny, nz = d->nz, np = d->npml;
    const int x = tl.y + threadIdx.x, y = tl.z + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;

    const long long sy = d->sy, sz = d->sz;
    const float *__restrict__ Ex = d->f[0];
    const float *__restrict__ Ey = d->f[1];
    const float *__restrict__ Ez = d->f[2];
    float *__restrict__ Hx = d->f[3];
    float *__restrict__ Hy = d->f[4];
    float *__restrict__ Hz = d->f[5];
    const float Db = d->Db, dtm0 = d->dtm0, dx = d->dx;

    const bool xin = x < nx - 1, yin = y < ny - 1;
    const bool hx_xy = x >= 1 && xin && yin;
    const bool hy_xy = y >= 1 && yin && xin;
    const bool hz_xy = yin && xin;
    const bool f_l = x < np - 1;
    const bool f_r = x >= nx - np && xin;
    const bool f_f = y < np - 1;
    const bool f_b = y >= ny - np && yin;
    const bool inject = d->mode == FDTD_MODE_MAIN && t < d->nt_inc;
    const int tz = d->tfsf_z;

    long long o = (long long)z0 * sz + (long long)y * sy + x;
    float ex_c = Ex[o], ey_c = Ey[o];
    for (int z = z0; z < z1; ++z, o += sz) {
        const bool zin = z < nz - 1;
        float hx = Hx[o], hy = Hy[o], hz = Hz[o];
        float ex_n = 0.f, ey_n = 0.f;
        if (zin) {
            ex_n = Ex[o + sz];
            ey_n = Ey[o + sz];
            const float ez_c = Ez[o];
            const float ez_yp = yin ? Ez[o + sy] : 0.f;
            const float ez_xp = xin ? Ez[o + 1] : 0.f;
            const float ey_xp = xin ? Ey[o + 1] : 0.f;
            const float ex_yp = yin ? Ex[o + sy] : 0.f;
            if (hx_xy) hx -= Db * ((ez_yp - ez_c) - (ey_n - ey_c));
            if (hy_xy) hy -= Db * ((ex_n - ex_c) - (ez_xp - ez_c));
            if (z >= 1 && hz_xy) hz -= Db * ((ey_xp - ey_c) - (ex_yp - ex_c));
            if (hz_xy) {
                if (z < np - 1) {                                   // down
                    const float bh = d->prof[BH_DLF][z], ch = d->prof[CH_DLF][z];
                    const long long p = (long long)z * d->du_sz + (long long)y * sy + x;
                    float *py = d->psi[pHyz_d], *px = d->psi[pHxz_d];
                    const float a = (bh * py[p]) + ((ch * (ex_n - ex_c)) / dx);
                    const float b = (bh * px[p]) + ((ch * (ey_n - ey_c)) / dx);
                    py[p] = a; px[p] = b;
                    hy -= dtm0 * a;
                    hx += dtm0 * b;
                }
                if (z >= nz - np) {                                 // up
                    const int zl = z - (nz - np);
                    const float bh = d->prof[BH_URB][zl], ch = d->prof[CH_URB][zl];
                    const long long p = (long long)zl * d->du_sz + (long long)y * sy + x;
                    float *py = d->psi[pHyz_u], *px = d->psi[pHxz_u];
                    const float a = (bh * py[p]) + ((ch * (ex_n - ex_c)) / dx);
                    const float b = (bh * px[p]) + ((ch * (ey_n - ey_c)) / dx);
                    py[p] = a; px[p] = b;
                    hy -= dtm0 * a;
                    hx += dtm0 * b;
                }
            }
            if (yin) {
                if (f_l) {                                          // left
                    const float bh = d->prof[BH_DLF][x], ch = d->prof[CH_DLF][x];
                    const long long p = (long long)z * d->lr_sz + (long long)y * np + x;
                    float *pz = d->psi[pHzx_l], *py = d->psi[pHyx_l];
                    const float a = (bh * pz[p]) + ((ch * (ey_xp - ey_c)) / dx);
                    const float b = (bh * py[p]) + ((ch * (ez_xp - ez_c)) / dx);
                    pz[p] = a; py[p] = b;
                    hz -= dtm0 * a;
                    hy += dtm0 * b;
                }
                if (f_r) {                                          // right
                    const int xl = x - (nx - np);
                    const float bh = d->prof[BH_URB][xl], ch = d->prof[CH_URB][xl];
                    const long long p = (long long)z * d->lr_sz + (long long)y * np + xl;
                    float *pz = d->psi[pHzx_r], *py = d->psi[pHyx_r];
                    const float a = (bh * pz[p]) + ((ch * (ey_xp - ey_c)) / dx);
                    const float b = (bh * py[p]) + ((ch * (ez_xp - ez_c)) / dx);
                    pz[p] = a; py[p] = b;
                    hz -= dtm0 * a;
                    hy += dtm0 * b;
                }
            }
            if (xin) {
                if (f_f) {                                          // front
                    const float bh = d->prof[BH_DLF][y], ch = d->prof[CH_DLF][y];
                    const long long p = (long long)z * d->fb_sz + (long long)y * sy + x;
                    float *pxy = d->psi[pHxy_f], *pzy = d->psi[pHzy_f];
                    const float a = (bh * pxy[p]) + ((ch * (ez_yp - ez_c)) / dx);
                    const float b = (bh * pzy[p]) + ((ch * (ex_yp - ex_c)) / dx);
                    pxy[p] = a; pzy[p] = b;
                    hx -= dtm0 * a;
                    hz += dtm0 * b;
                }
                if (f_b) {                                          // back
                    const int yl = y - (ny - np);
                    const float bh = d->prof[BH_URB][yl], ch = d->prof[CH_URB][yl];
                    const long long p = (long long)z * d->fb_sz + (long long)yl * sy + x;
                    float *pxy = d->psi[pHxy_b], *pzy = d->psi[pHzy_b];
                    const float a = (bh * pxy[p]) + ((ch * (ez_yp - ez_c)) / dx);
                    const float b = (bh * pzy[p]) + ((ch * (ex_yp - ex_c)) / dx);
                    pxy[p] = a; pzy[p] = b;
                    hx -= dtm0 * a;
                    hz += dtm0 * b;
                }
            }
        }
        if (inject && z == tz) {
            const long long q = ((long long)t * ny + y) * nx + x;
            hx += Db * d->beam[1 * d->beam_cs + q];   // ey_inc
            hy -= Db * d->beam[0 * d->beam_cs + q];   // ex_inc
        }
        Hx[o] = hx;
        Hy[o] = hy;
        Hz[o] = hz;
        ex_c = ex_n;
        ey_c = ey_n;
    }
}

// ---------------------------------------------------------------------------
// E half-step: Yee E update with per-cell Ca, Cb (cpu/field_updates.py:18-63),
// the six E-CPML corrections d,u,l,r,f,b (:66-191) and either the H_inc
// injection at tfsf_z (MAIN, cpu/main_sim.py:158-161) or the hard source at
// the feed-in plane (INCIDENT, cpu/incident_sim.py:142-144).  hx(z-1),
// hy(z-1) carry in registers up the column.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_e_half(const Desc *__restrict__ descs, const int4 *__restrict__ tiles, int ntiles) {
    const int4 tl = tiles[xcd_tile(blockIdx.x, ntiles)];
    const Desc *__restrict__ d = descs + tl.x;
    const int t = load_step(d->step);
    if (t >= d->nt) return;
    const int nx = d->nx, ny = d->ny, nz = d->nz, np = d->npml;
    const int x = tl.y + threadIdx.x, y = tl.z + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const int z0 = tl.w & 0xffff, z1 = tl.w >> 16;

    const long long sy = d->sy, sz = d->sz;
    float *__restrict__ Ex = d->f[0];
    float *__restrict__ Ey = d->f[1];
    float *__restrict__ Ez = d->f[2];
    const float *__restrict__ Hx = d->f[3];
    const float *__restrict__ Hy = d->f[4];
    const float *__restrict__ Hz = d->f[5];
    const float *__restrict__ Ca = d->ca;
    const float *__restrict__ Cb = d->cb;
    const int uni_z = d->uni_z;
    const float ca_u = d->ca_u, cb_u = d->cb_u, dx = d->dx;

    const bool xlo = x >= 1, ylo = y >= 1;
    const bool xin = x < nx - 1, yin = y < ny - 1;
    const bool ex_xy = ylo && yin && xin;
    const bool ey_xy = yin && xlo && xin;
    const bool ez_xy = ylo && yin && xlo && xin;
    const bool f_l = xlo && x < np;
    const bool f_r = x >= nx - np && xin;
    const bool f_f = ylo && y < np;
    const bool f_b = y >= ny - np && yin;
    const bool main_mode = d->mode == FDTD_MODE_MAIN;
    const bool inject = main_mode && t < d->nt_inc;
    const bool hard = !main_mode && t < d->n_src && ylo && yin && xin;
    const int tz = d->tfsf_z;

    long long o = (long long)z0 * sz + (long long)y * sy + x;
    float hx_m = 0.f, hy_m = 0.f;
    if (z0 >= 1) {
        hx_m = Hx[o - sz];
        hy_m = Hy[o - sz];
    }
    for (int z = z0; z < z1; ++z, o += sz) {
        const float hx_c = Hx[o], hy_c = Hy[o];
        float ex = Ex[o], ey = Ey[o], ez = Ez[o];
        if (z < nz - 1) {
            const float hz_c = Hz[o];
            const float hz_ym = ylo ? Hz[o - sy] : 0.f;
            const float hz_xm = xlo ? Hz[o - 1] : 0.f;
            const float hx_ym = ylo ? Hx[o - sy] : 0.f;
            const float hy_xm = xlo ? Hy[o - 1] : 0.f;
            const bool uni = z >= uni_z;                  // uniform (e.g. vacuum) plane
            const float ca = uni ? ca_u : Ca[o];
            const float cb = uni ? cb_u : Cb[o];
            const bool zlo = z >= 1;
            if (zlo && ex_xy) ex = (ex * ca) + (cb * ((hz_c - hz_ym) - (hy_c - hy_m)));
            if (zlo && ey_xy) ey = (ey * ca) + (cb * ((hx_c - hx_m) - (hz_c - hz_xm)));
            if (ez_xy) ez = (ez * ca) + (cb * ((hy_c - hy_xm) - (hx_c - hx_ym)));
            const float cbdx = cb * dx;
            if (yin && xin) {
                if (zlo && z < np) {                                // down
                    const float be = d->prof[BE_DLF][z], ce = d->prof[CE_DLF][z];
                    const long long p = (long long)z * d->du_sz + (long long)y * sy + x;
                    float *pey = d->psi[pEyz_d], *pex = d->psi[pExz_d];
                    const float a = (be * pey[p]) + ((ce * (hx_c - hx_m)) / dx);
                    const float b = (be * pex[p]) + ((ce * (hy_c - hy_m)) / dx);
                    pey[p] = a; pex[p] = b;
                    ey += cbdx * a;
                    ex -= cbdx * b;
                }
                if (z >= nz - np) {                                 // up
                    const int zl = z - (nz - np);
                    const float be = d->prof[BE_URB][zl], ce = d->prof[CE_URB][zl];
                    const long long p = (long long)zl * d->du_sz + (long long)y * sy + x;
                    float *pey = d->psi[pEyz_u], *pex = d->psi[pExz_u];
                    const float a = (be * pey[p]) + ((ce * (hx_c - hx_m)) / dx);
                    const float b = (be * pex[p]) + ((ce * (hy_c - hy_m)) / dx);
                    pey[p] = a; pex[p] = b;
                    ey += cbdx * a;
                    ex -= cbdx * b;
                }
            }
            if (yin) {
                if (f_l) {                                          // left
                    const float be = d->prof[BE_DLF][x], ce = d->prof[CE_DLF][x];
                    const long long p = (long long)z * d->lr_sz + (long long)y * np + x;
                    float *pez = d->psi[pEzx_l], *pey = d->psi[pEyx_l];
                    const float a = (be * pez[p]) + ((ce * (hy_c - hy_xm)) / dx);
                    const float b = (be * pey[p]) + ((ce * (hz_c - hz_xm)) / dx);
                    pez[p] = a; pey[p] = b;
                    ez += cbdx * a;
                    ey -= cbdx * b;
                }
                if (f_r) {                                          // right
                    const int xl = x - (nx - np);
                    const float be = d->prof[BE_URB][xl], ce = d->prof[CE_URB][xl];
                    const long long p = (long long)z * d->lr_sz + (long long)y * np + xl;
                    float *pez = d->psi[pEzx_r], *pey = d->psi[pEyx_r];
                    const float a = (be * pez[p]) + ((ce * (hy_c - hy_xm)) / dx);
                    const float b = (be * pey[p]) + ((ce * (hz_c - hz_xm)) / dx);
                    pez[p] = a; pey[p] = b;
                    ez += cbdx * a;
                    ey -= cbdx * b;
                }
            }
            if (xin) {
                if (f_f) {                                          // front
                    const float be = d->prof[BE_DLF][y], ce = d->prof[CE_DLF][y];
                    const long long p = (long long)z * d->fb_sz + (long long)y * sy + x;
                    float *pexy = d->psi[pExy_f], *pezy = d->psi[pEzy_f];
                    const float a = (be * pexy[p]) + ((ce * (hz_c - hz_ym)) / dx);
                    const float b = (be * pezy[p]) + ((ce * (hx_c - hx_ym)) / dx);
                    pexy[p] = a; pezy[p] = b;
                    ex += cbdx * a;
                    ez -= cbdx * b;
                }
                if (f_b) {                                          // back
                    const int yl = y - (ny - np);
                    const float be = d->prof[BE_URB][yl], ce = d->prof[CE_URB][yl];
                    const long long p = (long long)z * d->fb_sz + (long long)yl * sy + x;
                    float *pexy = d->psi[pExy_b], *pezy = d->psi[pEzy_b];
                    const float a = (be * pexy[p]) + ((ce * (hz_c - hz_ym)) / dx);
                    const float b = (be * pezy[p]) + ((ce * (hx_c - hx_ym)) / dx);
                    pexy[p] = a; pezy[p] = b;
                    ex += cbdx * a;
                    ez -= cbdx * b;
                }
            }
        }
        if (z == tz) {
            if (inject) {
                const long long q = ((long long)t * ny + y) * nx + x;
                ey += d->dte0dx * d->beam[2 * d->beam_cs + q];   // hx_inc
                ex -= d->dte0dx * d->beam[3 * d->beam_cs + q];   // hy_inc
            } else if (hard) {
                const long long q = ((long long)t * ny + y) * nx + x;
                ex = d->src[q];
                ey = d->src[d->src_cs + q];
            }
        }
        Ex[o] = ex;
        Ey[o] = ey;
        Ez[o] = ez;
        hx_m = hx_c;
        hy_m = hy_c;
    }
}

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
// The last block to finish advances every member's step counter, after all
// blocks have read it; the next launch in the stream sees the new value.
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
k_dft(const Desc *__restrict__ descs, int nmodels, unsigned int *done) {
    const Desc *__restrict__ d = descs + blockIdx.y;
    const int t = load_step(d->step);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < d->nt && i < d->dft_items) {
        const long long nsamp = d->face_off[d->nfaces];
        const long long sy = d->sy, sz = d->sz;
        const float *__restrict__ Ex = d->f[0];
        const float *__restrict__ Ey = d->f[1];
        const float *__restrict__ Ez = d->f[2];
        const float *__restrict__ Hx = d->f[3];
        const float *__restrict__ Hy = d->f[4];
        const float *__restrict__ Hz = d->f[5];
        if (i < nsamp) {
            int f = 0;
            while (f + 1 < d->nfaces && i >= d->face_off[f + 1]) ++f;
            const long long j = i - d->face_off[f];
            const int nb = d->face_nb[f];
            const int a = (int)(j / nb), b = (int)(j % nb);
            if (a >= 1 && b >= 1) {
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
            }
        } else if (d->rec) {
            // INCIDENT: record the read-out plane of ex, ey, hx, hy at step t
            // (cpu/incident_sim.py:153-156).
            const long long k = i - nsamp;
            const long long plane = (long long)d->ny * d->nx;
            const int c = (int)(k / plane);
            const long long r = k % plane;
            const int yy = (int)(r / d->nx), xx = (int)(r % d->nx);
            const int src_comp[4] = {0, 1, 3, 4};
            const float v = d->f[src_comp[c]][(long long)d->read_out * sz + (long long)yy * sy + xx];
            d->rec[c * d->beam_cs + (long long)t * plane + r] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned int total = gridDim.x * gridDim.y;
        const unsigned int prev = atomicAdd(done, 1u);
        if (prev == total - 1) {
            for (int m = 0; m < nmodels; ++m) {
                int *sp = descs[m].step;
                if (load_step(sp) < descs[m].nt) atomicAdd(sp, 1);
            }
            atomicExch(done, 0u);
            __threadfence();
        }
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
