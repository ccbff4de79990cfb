// Material models built on the device: per-cell (eps_r, sigma) from
// layers below int32 surface maps, voxel inclusions and cylinders, then
// the Yee E-update coefficients Ca, Cb (materials/materials.py:23-68)
// written straight into a simulation's pitched material arrays.
//
// Arithmetic follows the host generators operation by operation, in IEEE
// double for (eps_r, sigma) and float32 for Ca, Cb, with no contraction.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdtd_kernels.h"

namespace fdtd {

// SplitMix64 finalizer (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014) as a counter-based hash.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Draw j (0..3) of cell `cell` in stream `seed`, uniform on [0, 1) with 53
// random bits.
__device__ __forceinline__ double draw01(uint32_t seed, uint64_t cell, int j) {
    const uint64_t key = ((uint64_t)seed << 36) ^ (cell << 2) ^ (uint64_t)j;
    return (double)(mix64(key) >> 11) * 0x1.0p-53;
}

// numpy.interp between two table points (f0, y0), (f1, y1) at f0 <= x <= f1:
// the end points exactly, else slope (x - f0) + y0.
__device__ __forceinline__ double interp2(double x, double f0, double f1, double y0, double y1) {
    if (x == f0) return y0;
    if (x == f1) return y1;
    const double slope = (y1 - y0) / (f1 - f0);
    return slope * (x - f0) + y0;
}

__device__ __forceinline__ double poly(const double (&p)[3], double mv) {
    return (p[0] + (p[1] * mv)) + (p[2] * (mv * mv));
}

// Cylinder voxelisation: block (c, k) tests part k of cylinder c's cell box
// and raises each inside cell's owner to c + 1 (the last cylinder wins, as
// when the host generator paints them in order).
__global__ void __launch_bounds__(256)
k_mat_cylinders(const FdtdMatCylinder *__restrict__ cyl, double dx, int *__restrict__ owner, long long sy,
                long long sz) {
    const int c = blockIdx.x;
    const FdtdMatCylinder C = cyl[c];
    const long long bz = C.hi[0] - C.lo[0], by = C.hi[1] - C.lo[1], bx = C.hi[2] - C.lo[2];
    const long long n = bz * by * bx;
    const double half_len = 0.5 * C.length, r2 = C.radius * C.radius;
    for (long long i = (long long)blockIdx.y * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.y * blockDim.x) {
        const long long x = C.lo[2] + i % bx, r = i / bx;
        const long long y = C.lo[1] + r % by, z = C.lo[0] + r / by;
        const double dz = (double)z * dx - C.c[0];
        const double dy = (double)y * dx - C.c[1];
        const double dxx = (double)x * dx - C.c[2];
        const double t = ((dz * C.u[0]) + (dy * C.u[1])) + (dxx * C.u[2]);
        const double perp2 = (((dz * dz) + (dy * dy)) + (dxx * dxx)) - (t * t);
        if (fabs(t) <= half_len && perp2 <= r2) atomicMax(&owner[z * sz + y * sy + x], c + 1);
    }
}

// One thread per cell (z, y, x) of the whole grid: eps_r, sigma, then Ca,
// Cb.  `owner` aliases Cb: each thread reads its own cell's entry before
// writing it.
__global__ void __launch_bounds__(256)
k_mat_build(FdtdMatSpec spec, const int32_t *__restrict__ surf, const FdtdMatCylinder *__restrict__ cyl,
            int nx, int ny, int nz, long long sy, long long sz, float *__restrict__ Ca, float *Cb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)nx * ny;
    if (i >= plane * nz) return;
    const int z = (int)(i / plane);
    const long long rem = i - (long long)z * plane;
    const int y = (int)(rem / nx), x = (int)(rem - (long long)y * nx);
    const long long o = z * sz + y * sy + x;
    float eps = 1.0f, sig = 0.0f;
    if (z < spec.nz_mat) {
        const uint64_t cell = (uint64_t)i;
        if (z < spec.vox_z1 && draw01(spec.vox_seed, cell, 3) < spec.vox_frac) {
            eps = (float)spec.vox_eps;
            sig = (float)spec.vox_sigma;
        }
        if (spec.ncyl > 0) {
            const int own = reinterpret_cast<const int *>(Cb)[o];
            if (own > 0) {
                eps = (float)cyl[own - 1].eps;
                sig = (float)cyl[own - 1].sigma;
            }
        }
        for (int k = 0; k < spec.nlayers; ++k) {
            const int h = surf[(long long)k * plane + rem];
            if (z > h) continue;
            const FdtdMatLayer &L = spec.layer[k];
            if (L.kind == FDTD_LAYER_CONST) {
                eps = (float)L.eps;
                sig = (float)L.sigma;
            } else if (L.kind == FDTD_LAYER_RANDOM) {
                const double u1 = 1.0 - draw01(L.seed, cell, 0), u2 = draw01(L.seed, cell, 1);
                const double n = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
                double e = L.eps + L.eps_std * n;
                if (e < 1.0) e = 1.0;
                eps = (float)e;
                sig = (float)(L.sigma + (L.sigma_hi - L.sigma) * draw01(L.seed, cell, 2));
            } else {
                const int h0 = surf[(long long)(k - 1) * plane + rem];
                const double mv = ((L.mv_bottom * (double)(h - z)) + (L.mv_top * (double)(z - h0))) /
                                  (double)(h - h0);
                const double re = interp2(L.freq, L.f0, L.f1, poly(L.poly[0][0], mv), poly(L.poly[1][0], mv));
                const double im = interp2(L.freq, L.f0, L.f1, poly(L.poly[0][1], mv), poly(L.poly[1][1], mv));
                eps = (float)re;
                // |imag(re - j im)| eps0 2 pi f
                sig = (float)(((fabs(im) * L.eps0) * 2.0) * M_PI * L.freq);
            }
            break;
        }
    }
    const float perm = eps * spec.eps0;
    const float loss = (sig * spec.dt) / (2.0f * perm);
    Ca[o] = (1.0f - loss) / (1.0f + loss);
    Cb[o] = (spec.dt / (perm * spec.dx)) / (1.0f + loss);
}

}  // namespace fdtd
