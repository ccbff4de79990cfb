// Incident-beam excitation synthesis on the device.
//
// Device form of IncidentBeam.create_inc_excitation
// (fdtdsarsim/incidentbeam.py:206-313) and taper_k (:316-433):
//   for every temporal-frequency bin k of the pulse spectrum
//     S_k(ky, kx)  tapered-wave spectrum (Braunisch), x and y components
//     A_k = fftshift(fft2(ifftshift(S_k)))          (FP64, hipFFT Z2Z)
//     pulse[k] = complex64(pulse_fft[k] * A_k)
//   pulse[:, y, x] = ifft over k                     (complex64, hipFFT C2C)
//   pulse *= 10 / max |pulse|                        (both components)
// and the imaginary parts of the first n_src planes become the hard source
// of the incident FDTD (cpu/incident_sim.py:110-111).
//
// ifftshift of a spectrum sampled on the fftshift'ed grid is the same
// spectrum sampled on the unshifted fftfreq grid, so S_k is evaluated there
// directly; the output fftshift is an index map when A_k is stored.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "fdtd_kernels.h"

namespace fdtd {

struct TaperSpec {
    int nx, ny;
    double val_x, val_y;            // 1 / (n * dx) of numpy's fftfreq
    double radius;                  // gaussian radius [m]
    double sin_th, sin_phi, cos_phi;
    double Eh_re, Eh_im, Ev_re, Ev_im;
};

struct cdbl { double re, im; };
__device__ __forceinline__ cdbl cadd(cdbl a, cdbl b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cdbl cmul(cdbl a, cdbl b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cdbl cscale(double s, cdbl a) { return {s * a.re, s * a.im}; }

__device__ __forceinline__ double fftfreq_scaled(int i, int n, double val) {
    const int s = (i < (n + 1) / 2) ? i : i - n;   // numpy fftfreq order
    return 2.0 * 3.141592653589793 * ((double)s * val);
}

// Spectra of bins [k0, k0 + nb) in unshifted FFT order: out_x/out_y [nb][ny][nx].
__global__ void __launch_bounds__(256)
k_taper_spectrum(TaperSpec sp, const double *__restrict__ all_k, int k0, int nb,
                 double2 *__restrict__ out_x, double2 *__restrict__ out_y) {
    const long long plane = (long long)sp.ny * sp.nx;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane * nb) return;
    const int b = (int)(idx / plane);
    const long long r = idx - (long long)b * plane;
    const int j = (int)(r / sp.nx), i = (int)(r - (long long)j * sp.nx);
    const double k = all_k[k0 + b];
    if (k == 0.0) {
        out_x[idx] = make_double2(0.0, 0.0);
        out_y[idx] = make_double2(0.0, 0.0);
        return;
    }
    const double kx = fftfreq_scaled(i, sp.nx, sp.val_x);
    const double ky = fftfreq_scaled(j, sp.ny, sp.val_y);
    const double kp = hypot(kx, ky);
    // k_z = conj(sqrt(complex64(k^2 - k_p^2))): the argument is real, so the
    // complex64 square root is a float32 square root on one axis
    const float arg = (float)(k * k - kp * kp);
    const cdbl kz = arg >= 0.f ? cdbl{(double)sqrtf(arg), 0.0} : cdbl{0.0, -(double)sqrtf(-arg)};
    // central plane wave of this bin (incidentbeam.py:351-356, 392-406); its
    // basis depends on the sign of k for negative-frequency bins
    const double k_ix = (k * sp.sin_th) * sp.cos_phi;
    const double k_iy = (k * sp.sin_th) * sp.sin_phi;
    const double k_ip = hypot(k_ix, k_iy);
    const double k_iz = sqrt(k * k - k_ip * k_ip);
    double hor_ix, hor_iy, ver_ix, ver_iy, ver_iz;
    if (k_ip > 0) {
        hor_ix = k_iy / k_ip;
        hor_iy = -k_ix / k_ip;
        ver_ix = (k_iz * k_ix) / (k * k_ip);
        ver_iy = (k_iz * k_iy) / (k * k_ip);
        ver_iz = k_ip / k;
    } else {
        hor_ix = sp.sin_phi;
        hor_iy = -sp.cos_phi;
        ver_ix = sp.cos_phi;
        ver_iy = sp.sin_phi;
        ver_iz = 0.0;
    }
    const double rad2 = sp.radius * sp.radius;
    const double dxk = kx - k_ix, dyk = ky - k_iy;
    const double gauss = (rad2 / (4 * 3.141592653589793)) * exp(-rad2 * ((dxk * dxk) + (dyk * dyk)) / 4);
    cdbl hor_x, hor_y, ver_x, ver_y, ver_z;
    if (kp > 0.0) {
        hor_x = {ky / kp, 0.0};
        hor_y = {-kx / kp, 0.0};
        const double den = k * kp;
        ver_x = {(kz.re * kx) / den, (kz.im * kx) / den};
        ver_y = {(kz.re * ky) / den, (kz.im * ky) / den};
        ver_z = {kp / k, 0.0};
    } else {
        hor_x = {sp.sin_phi, 0.0};
        hor_y = {-sp.cos_phi, 0.0};
        ver_x = {sp.cos_phi, 0.0};
        ver_y = {sp.sin_phi, 0.0};
        ver_z = {0.0, 0.0};
    }
    const cdbl Eh = {sp.Eh_re, sp.Eh_im}, Ev = {sp.Ev_re, sp.Ev_im};
    const cdbl eh = cadd(cmul(Eh, cadd(cscale(hor_ix, hor_x), cscale(hor_iy, hor_y))),
                         cmul(Ev, cadd(cscale(ver_ix, hor_x), cscale(ver_iy, hor_y))));
    const cdbl ev = cadd(cmul(Eh, cadd(cscale(hor_ix, ver_x), cscale(hor_iy, ver_y))),
                         cmul(Ev, cadd(cadd(cscale(ver_ix, ver_x), cscale(ver_iy, ver_y)),
                                       cscale(ver_iz, ver_z))));
    const cdbl xc = cscale(gauss, cadd(cmul(eh, hor_x), cmul(ev, ver_x)));
    const cdbl yc = cscale(gauss, cadd(cmul(eh, hor_y), cmul(ev, ver_y)));
    out_x[idx] = make_double2(xc.re, xc.im);
    out_y[idx] = make_double2(yc.re, yc.im);
}

// pulse[k0 + b][y][x] = complex64(pulse_fft[k0 + b] * A[b][(y + (ny+1)/2) % ny][(x + (nx+1)/2) % nx])
__global__ void __launch_bounds__(256)
k_shift_scale(int nx, int ny, int k0, int nb, const double2 *__restrict__ pulse_fft,
              const double2 *__restrict__ ax, const double2 *__restrict__ ay,
              float2 *__restrict__ px, float2 *__restrict__ py) {
    const long long plane = (long long)ny * nx;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= plane * nb) return;
    const int b = (int)(idx / plane);
    const long long r = idx - (long long)b * plane;
    const int y = (int)(r / nx), x = (int)(r - (long long)y * nx);
    const int ys = (y + (ny + 1) / 2) % ny, xs = (x + (nx + 1) / 2) % nx;
    const long long src = (long long)b * plane + (long long)ys * nx + xs;
    const long long dst = (long long)(k0 + b) * plane + r;
    const double2 w = pulse_fft[k0 + b];
    const double2 a = ax[src], c = ay[src];
    px[dst] = make_float2((float)(w.x * a.x - w.y * a.y), (float)(w.x * a.y + w.y * a.x));
    py[dst] = make_float2((float)(w.x * c.x - w.y * c.y), (float)(w.x * c.y + w.y * c.x));
}

// out = scale * v (float32), the 1/n of the inverse transform
__global__ void __launch_bounds__(256)
k_scale_c64(float2 *__restrict__ v, long long n, float scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 a = v[i];
        v[i] = make_float2(a.x * scale, a.y * scale);
    }
}

// max |v| over both arrays; non-negative floats order like their bit patterns
__global__ void __launch_bounds__(256)
k_max_abs(const float2 *__restrict__ a, const float2 *__restrict__ b, long long n, unsigned int *__restrict__ out) {
    float m = 0.f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        m = fmaxf(m, hypotf(a[i].x, a[i].y));
        m = fmaxf(m, hypotf(b[i].x, b[i].y));
    }
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, kWave));
    if ((threadIdx.x % kWave) == 0) atomicMax(out, __float_as_uint(m));
}

// src[c][t][y][x] = imag(complex64(complex128(pulse) * (10 / max)))
__global__ void __launch_bounds__(256)
k_source_planes(const float2 *__restrict__ px, const float2 *__restrict__ py, long long n,
                const unsigned int *__restrict__ maxbits, float *__restrict__ src) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = 10.0 / (double)__uint_as_float(*maxbits);
    src[i] = (float)((double)px[i].y * s);
    src[n + i] = (float)((double)py[i].y * s);
}

}  // namespace fdtd

using namespace fdtd;

// Synthesise the excitation on the device; returns the hard-source planes
// [2][n_src][ny][nx] (device, caller frees) via *src_out.
int beam_synthesize_source(const double *pulse_fft, const double *all_k, int nt_inc, int n_src, int nx, int ny,
                           double dx, double theta_deg, double phi_deg, double radius, int pol, hipStream_t st,
                           float **src_out, char *err, size_t errlen) {
    auto fail = [&](const char *what, int code) {
        snprintf(err, errlen, "%s (%d)", what, code);
        return -2;
    };
#define HT(expr)                                        \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return fail(#expr, (int)e_); \
    } while (0)
#define FT(expr)                                          \
    do {                                                  \
        hipfftResult r_ = (expr);                         \
        if (r_ != HIPFFT_SUCCESS) return fail(#expr, (int)r_); \
    } while (0)
    const double pi = 3.141592653589793;
    TaperSpec sp{};
    sp.nx = nx; sp.ny = ny;
    sp.val_x = 1.0 / (nx * dx);
    sp.val_y = 1.0 / (ny * dx);
    sp.radius = radius;
    const double th = theta_deg * (pi / 180.0), ph = phi_deg * (pi / 180.0);
    sp.sin_th = std::sin(th);
    sp.sin_phi = std::sin(ph);
    sp.cos_phi = std::cos(ph);
    double Eh[2] = {0, 0}, Ev[2] = {0, 0};
    switch (pol) {
    case 0: Eh[0] = 1; break;                    // HH
    case 1: Ev[0] = 1; break;                    // VV
    case 2: Eh[1] = -1; Ev[0] = 1; break;        // RHCP
    case 3: Eh[1] = 1; Ev[0] = 1; break;         // LHCP
    default: return fail("unknown polarisation", pol);
    }
    sp.Eh_re = Eh[0]; sp.Eh_im = Eh[1]; sp.Ev_re = Ev[0]; sp.Ev_im = Ev[1];

    const long long plane = (long long)nx * ny;
    const long long total = plane * nt_inc;
    double *d_k = nullptr;
    double2 *d_pf = nullptr, *d_ax = nullptr, *d_ay = nullptr;
    float2 *d_px = nullptr, *d_py = nullptr;
    unsigned int *d_max = nullptr;
    float *d_src = nullptr;
    const int nb = (int)std::max(1LL, std::min<long long>(nt_inc, (256LL << 20) / (plane * 16)));
    HT(hipMalloc(&d_k, nt_inc * sizeof(double)));
    HT(hipMalloc(&d_pf, nt_inc * sizeof(double2)));
    HT(hipMalloc(&d_ax, (size_t)nb * plane * sizeof(double2)));
    HT(hipMalloc(&d_ay, (size_t)nb * plane * sizeof(double2)));
    HT(hipMalloc(&d_px, (size_t)total * sizeof(float2)));
    HT(hipMalloc(&d_py, (size_t)total * sizeof(float2)));
    HT(hipMalloc(&d_max, sizeof(unsigned int)));
    HT(hipMemcpyAsync(d_k, all_k, nt_inc * sizeof(double), hipMemcpyHostToDevice, st));
    HT(hipMemcpyAsync(d_pf, pulse_fft, nt_inc * sizeof(double2), hipMemcpyHostToDevice, st));

    // 2-D forward transforms of nb bins at a time (FP64)
    int dims[2] = {ny, nx};
    hipfftHandle p2 = 0, p1 = 0;
    FT(hipfftPlanMany(&p2, 2, dims, nullptr, 1, (int)plane, nullptr, 1, (int)plane, HIPFFT_Z2Z, nb));
    FT(hipfftSetStream(p2, st));
    for (int k0 = 0; k0 < nt_inc; k0 += nb) {
        const int cnt = std::min(nb, nt_inc - k0);
        const long long n = plane * cnt;
        const unsigned g = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_taper_spectrum, dim3(g), dim3(256), 0, st, sp, d_k, k0, cnt, d_ax, d_ay);
        HT(hipGetLastError());
        if (cnt == nb) {
            FT(hipfftExecZ2Z(p2, (hipfftDoubleComplex *)d_ax, (hipfftDoubleComplex *)d_ax, HIPFFT_FORWARD));
            FT(hipfftExecZ2Z(p2, (hipfftDoubleComplex *)d_ay, (hipfftDoubleComplex *)d_ay, HIPFFT_FORWARD));
        } else {
            hipfftHandle pr = 0;
            FT(hipfftPlanMany(&pr, 2, dims, nullptr, 1, (int)plane, nullptr, 1, (int)plane, HIPFFT_Z2Z, cnt));
            FT(hipfftSetStream(pr, st));
            FT(hipfftExecZ2Z(pr, (hipfftDoubleComplex *)d_ax, (hipfftDoubleComplex *)d_ax, HIPFFT_FORWARD));
            FT(hipfftExecZ2Z(pr, (hipfftDoubleComplex *)d_ay, (hipfftDoubleComplex *)d_ay, HIPFFT_FORWARD));
            HT(hipStreamSynchronize(st));
            FT(hipfftDestroy(pr));
        }
        hipLaunchKernelGGL(k_shift_scale, dim3(g), dim3(256), 0, st, nx, ny, k0, cnt, d_pf, d_ax, d_ay, d_px, d_py);
        HT(hipGetLastError());
    }
    HT(hipStreamSynchronize(st));
    FT(hipfftDestroy(p2));
    (void)hipFree(d_ax);
    (void)hipFree(d_ay);
    d_ax = d_ay = nullptr;

    // inverse transform along t for every (y, x) (complex64, scaled by 1/n)
    int n1[1] = {nt_inc};
    FT(hipfftPlanMany(&p1, 1, n1, n1, (int)plane, 1, n1, (int)plane, 1, HIPFFT_C2C, (int)plane));
    FT(hipfftSetStream(p1, st));
    FT(hipfftExecC2C(p1, (hipfftComplex *)d_px, (hipfftComplex *)d_px, HIPFFT_BACKWARD));
    FT(hipfftExecC2C(p1, (hipfftComplex *)d_py, (hipfftComplex *)d_py, HIPFFT_BACKWARD));
    const unsigned gt = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_scale_c64, dim3(gt), dim3(256), 0, st, d_px, total, 1.0f / (float)nt_inc);
    hipLaunchKernelGGL(k_scale_c64, dim3(gt), dim3(256), 0, st, d_py, total, 1.0f / (float)nt_inc);
    HT(hipMemsetAsync(d_max, 0, sizeof(unsigned int), st));
    hipLaunchKernelGGL(k_max_abs, dim3(2048), dim3(256), 0, st, d_px, d_py, total, d_max);
    const long long ns = plane * n_src;
    HT(hipMalloc(&d_src, std::max<long long>(1, 2 * ns) * sizeof(float)));
    if (ns > 0)
        hipLaunchKernelGGL(k_source_planes, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, d_px, d_py, ns,
                           d_max, d_src);
    HT(hipGetLastError());
    HT(hipStreamSynchronize(st));
    FT(hipfftDestroy(p1));
    (void)hipFree(d_k);
    (void)hipFree(d_pf);
    (void)hipFree(d_px);
    (void)hipFree(d_py);
    (void)hipFree(d_max);
    *src_out = d_src;
    return 0;
#undef HT
#undef FT
}
