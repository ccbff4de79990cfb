/*
 * ORACLE — test infrastructure only.  This file is the CPU restatement of
 * the reference's FDTD loops that the parity tests, the smoke check and the
 * CPU baseline compare the HIP library against.  Nothing in the product
 * links or calls it.
 *
 * Instantiated twice by fdtd_oracle.c with T = float (suffix f32) and
 * T = double (suffix f64); build with -ffp-contract=off so every expression
 * rounds once per operation, as numba does without fastmath.
 *
 * Each pass follows one numba kernel of fdtdsarsim/cpu/field_updates.py
 * (cited per function) with the same index ranges and the same order of
 * floating-point operations.  The time loop follows cpu/main_sim.py:140-195
 * (main) and cpu/incident_sim.py:134-161 (incident); CPML coefficients are
 * the 1-D profiles the reference tiles to 3-D (absorbers.py:123-185), read
 * at the index the tiled array would have.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* 3-D index into a (nz, ny, nx) C-order array. */
#define I3(z, y, x) (((size_t)(z) * (size_t)ny + (size_t)(y)) * (size_t)nx + (size_t)(x))

typedef struct {
    T *ex, *ey, *ez, *hx, *hy, *hz;          /* (nz, ny, nx) */
    /* E psi: d/u (npml, ny, nx); l/r (nz, ny, npml); f/b (nz, npml, nx) */
    T *psiExz_d, *psiEyz_d, *psiExz_u, *psiEyz_u;
    T *psiEyx_l, *psiEzx_l, *psiEyx_r, *psiEzx_r;
    T *psiExy_f, *psiEzy_f, *psiExy_b, *psiEzy_b;
    T *psiHxz_d, *psiHyz_d, *psiHxz_u, *psiHyz_u;
    T *psiHyx_l, *psiHzx_l, *psiHyx_r, *psiHzx_r;
    T *psiHxy_f, *psiHzy_f, *psiHxy_b, *psiHzy_b;
    /* phasors, interleaved (re, im): z faces (ny_ph, nx_ph), x faces
     * (nz_ph, ny_ph), y faces (nz_ph, nx_ph); order of the reference's
     * save list (cpu/main_sim.py:220-229). */
    T *ph[24];
} FN(OracleState_);

typedef struct {
    int nx, ny, nz, npml;
    int tfsf_z;            /* main: injection plane; incident: feed_in */
    int read_out;          /* incident only */
    int nt_inc;            /* main: injected steps; incident: nt_pulse */
    T Db, dtm0, dx, dte0dx;
    double omega, dt;
    const T *Ca, *Cb;      /* (nz, ny, nx) */
    const T *prof;         /* be_dlf ce_dlf bh_dlf ch_dlf be_urb ce_urb bh_urb ch_urb, npml each */
    const T *inc;          /* main: ex,ey,hx,hy each (nt_inc, ny, nx); incident: src ex,ey imag (nt_pulse, ny, nx) */
    T *rec;                /* incident: recorded ex,ey,hx,hy each (n_rec, ny, nx) */
    int nthreads;
} FN(OracleGeom_);

#define ST FN(OracleState_)
#define GE FN(OracleGeom_)

/* ---------------- E Yee updates: cpu/field_updates.py:18-63 ---------------- */
static void FN(ex_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx;
    T *ex = s->ex; const T *hy = s->hy, *hz = s->hz, *ca = g->Ca, *cb = g->Cb;
    for (int z = z0; z < z1; ++z)
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T scaled = ex[I3(z, y, x)] * ca[I3(z, y, x)];
                T dhz = hz[I3(z, y, x)] - hz[I3(z, y - 1, x)];
                T dhy = hy[I3(z, y, x)] - hy[I3(z - 1, y, x)];
                ex[I3(z, y, x)] = scaled + (cb[I3(z, y, x)] * (dhz - dhy));
            }
}
static void FN(ey_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx;
    T *ey = s->ey; const T *hx = s->hx, *hz = s->hz, *ca = g->Ca, *cb = g->Cb;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 1; x < nx - 1; ++x) {
                T scaled = ey[I3(z, y, x)] * ca[I3(z, y, x)];
                T dhx = hx[I3(z, y, x)] - hx[I3(z - 1, y, x)];
                T dhz = hz[I3(z, y, x)] - hz[I3(z, y, x - 1)];
                ey[I3(z, y, x)] = scaled + (cb[I3(z, y, x)] * (dhx - dhz));
            }
}
static void FN(ez_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx;
    T *ez = s->ez; const T *hx = s->hx, *hy = s->hy, *ca = g->Ca, *cb = g->Cb;
    for (int z = z0; z < z1; ++z)
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 1; x < nx - 1; ++x) {
                T scaled = ez[I3(z, y, x)] * ca[I3(z, y, x)];
                T dhx = hx[I3(z, y, x)] - hx[I3(z, y - 1, x)];
                T dhy = hy[I3(z, y, x)] - hy[I3(z, y, x - 1)];
                ez[I3(z, y, x)] = scaled + (cb[I3(z, y, x)] * (dhy - dhx));
            }
}

/* ---------------- H Yee updates: cpu/field_updates.py:198-240 -------------- */
static void FN(hx_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx; const T db = g->Db;
    T *hx = s->hx; const T *ey = s->ey, *ez = s->ez;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 1; x < nx - 1; ++x) {
                T dez = ez[I3(z, y + 1, x)] - ez[I3(z, y, x)];
                T dey = ey[I3(z + 1, y, x)] - ey[I3(z, y, x)];
                hx[I3(z, y, x)] -= db * (dez - dey);
            }
}
static void FN(hy_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx; const T db = g->Db;
    T *hy = s->hy; const T *ex = s->ex, *ez = s->ez;
    for (int z = z0; z < z1; ++z)
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T dex = ex[I3(z + 1, y, x)] - ex[I3(z, y, x)];
                T dez = ez[I3(z, y, x + 1)] - ez[I3(z, y, x)];
                hy[I3(z, y, x)] -= db * (dex - dez);
            }
}
static void FN(hz_update_)(const GE *g, ST *s, int z0, int z1) {
    const int ny = g->ny, nx = g->nx; const T db = g->Db;
    T *hz = s->hz; const T *ex = s->ex, *ey = s->ey;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T dex = ex[I3(z, y + 1, x)] - ex[I3(z, y, x)];
                T dey = ey[I3(z, y, x + 1)] - ey[I3(z, y, x)];
                hz[I3(z, y, x)] -= db * (dey - dex);
            }
}

/* Profile accessors: be/ce/bh/ch for the d,l,f (dlf) and u,r,b (urb) faces. */
#define P_BE_DLF(i) (g->prof[0 * np_ + (i)])
#define P_CE_DLF(i) (g->prof[1 * np_ + (i)])
#define P_BH_DLF(i) (g->prof[2 * np_ + (i)])
#define P_CH_DLF(i) (g->prof[3 * np_ + (i)])
#define P_BE_URB(i) (g->prof[4 * np_ + (i)])
#define P_CE_URB(i) (g->prof[5 * np_ + (i)])
#define P_BH_URB(i) (g->prof[6 * np_ + (i)])
#define P_CH_URB(i) (g->prof[7 * np_ + (i)])

/* psi index helpers: (npml, ny, nx), (nz, ny, npml), (nz, npml, nx) */
#define IZF(z, y, x) (((size_t)(z) * (size_t)ny + (size_t)(y)) * (size_t)nx + (size_t)(x))
#define IXF(z, y, x) (((size_t)(z) * (size_t)ny + (size_t)(y)) * (size_t)np_ + (size_t)(x))
#define IYF(z, y, x) (((size_t)(z) * (size_t)np_ + (size_t)(y)) * (size_t)nx + (size_t)(x))

/* ---------------- E CPML: cpu/field_updates.py:66-191 ---------------------- */
static void FN(cpml_e_down_)(const GE *g, ST *s) {          /* :66-83 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx;
    for (int z = 1; z < np_; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T be = P_BE_DLF(z), ce = P_CE_DLF(z), cb = g->Cb[I3(z, y, x)];
                size_t p = IZF(z, y, x);
                s->psiEyz_d[p] = (be * s->psiEyz_d[p]) + (ce * (s->hx[I3(z, y, x)] - s->hx[I3(z - 1, y, x)]) / dx);
                s->psiExz_d[p] = (be * s->psiExz_d[p]) + (ce * (s->hy[I3(z, y, x)] - s->hy[I3(z - 1, y, x)]) / dx);
                s->ey[I3(z, y, x)] += cb * dx * s->psiEyz_d[p];
                s->ex[I3(z, y, x)] -= cb * dx * s->psiExz_d[p];
            }
}
static void FN(cpml_e_up_)(const GE *g, ST *s) {            /* :86-105 */
    const int ny = g->ny, nx = g->nx, nz = g->nz, np_ = g->npml; const T dx = g->dx;
    for (int z = 0; z < np_ - 1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                int zf = nz - np_ + z;
                T be = P_BE_URB(z), ce = P_CE_URB(z), cb = g->Cb[I3(zf, y, x)];
                size_t p = IZF(z, y, x);
                s->psiEyz_u[p] = be * s->psiEyz_u[p] + (ce * (s->hx[I3(zf, y, x)] - s->hx[I3(zf - 1, y, x)]) / dx);
                s->psiExz_u[p] = be * s->psiExz_u[p] + (ce * (s->hy[I3(zf, y, x)] - s->hy[I3(zf - 1, y, x)]) / dx);
                s->ey[I3(zf, y, x)] += cb * dx * s->psiEyz_u[p];
                s->ex[I3(zf, y, x)] -= cb * dx * s->psiExz_u[p];
            }
}
static void FN(cpml_e_left_)(const GE *g, ST *s, int z0, int z1) {   /* :108-125 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 1; x < np_; ++x) {
                T be = P_BE_DLF(x), ce = P_CE_DLF(x), cb = g->Cb[I3(z, y, x)];
                size_t p = IXF(z, y, x);
                s->psiEzx_l[p] = (be * s->psiEzx_l[p]) + (ce * (s->hy[I3(z, y, x)] - s->hy[I3(z, y, x - 1)]) / dx);
                s->psiEyx_l[p] = (be * s->psiEyx_l[p]) + (ce * (s->hz[I3(z, y, x)] - s->hz[I3(z, y, x - 1)]) / dx);
                s->ez[I3(z, y, x)] += cb * dx * s->psiEzx_l[p];
                s->ey[I3(z, y, x)] -= cb * dx * s->psiEyx_l[p];
            }
}
static void FN(cpml_e_right_)(const GE *g, ST *s, int z0, int z1) {  /* :128-148 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < np_ - 1; ++x) {
                int xf = nx - np_ + x;
                T be = P_BE_URB(x), ce = P_CE_URB(x), cb = g->Cb[I3(z, y, xf)];
                size_t p = IXF(z, y, x);
                s->psiEzx_r[p] = (be * s->psiEzx_r[p]) + (ce * (s->hy[I3(z, y, xf)] - s->hy[I3(z, y, xf - 1)]) / dx);
                s->psiEyx_r[p] = (be * s->psiEyx_r[p]) + (ce * (s->hz[I3(z, y, xf)] - s->hz[I3(z, y, xf - 1)]) / dx);
                s->ez[I3(z, y, xf)] += cb * dx * s->psiEzx_r[p];
                s->ey[I3(z, y, xf)] -= cb * dx * s->psiEyx_r[p];
            }
}
static void FN(cpml_e_front_)(const GE *g, ST *s, int z0, int z1) {  /* :151-169 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx;
    for (int z = z0; z < z1; ++z)
        for (int y = 1; y < np_; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T be = P_BE_DLF(y), ce = P_CE_DLF(y), cb = g->Cb[I3(z, y, x)];
                size_t p = IYF(z, y, x);
                s->psiExy_f[p] = (be * s->psiExy_f[p]) + (ce * (s->hz[I3(z, y, x)] - s->hz[I3(z, y - 1, x)]) / dx);
                s->psiEzy_f[p] = (be * s->psiEzy_f[p]) + (ce * (s->hx[I3(z, y, x)] - s->hx[I3(z, y - 1, x)]) / dx);
                s->ex[I3(z, y, x)] += cb * dx * s->psiExy_f[p];
                s->ez[I3(z, y, x)] -= cb * dx * s->psiEzy_f[p];
            }
}
static void FN(cpml_e_back_)(const GE *g, ST *s, int z0, int z1) {   /* :172-191 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < np_ - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                int yf = ny - np_ + y;
                T be = P_BE_URB(y), ce = P_CE_URB(y), cb = g->Cb[I3(z, yf, x)];
                size_t p = IYF(z, y, x);
                s->psiExy_b[p] = (be * s->psiExy_b[p]) + (ce * (s->hz[I3(z, yf, x)] - s->hz[I3(z, yf - 1, x)]) / dx);
                s->psiEzy_b[p] = (be * s->psiEzy_b[p]) + (ce * (s->hx[I3(z, yf, x)] - s->hx[I3(z, yf - 1, x)]) / dx);
                s->ex[I3(z, yf, x)] += cb * dx * s->psiExy_b[p];
                s->ez[I3(z, yf, x)] -= cb * dx * s->psiEzy_b[p];
            }
}

/* ---------------- H CPML: cpu/field_updates.py:243-366 --------------------- */
static void FN(cpml_h_down_)(const GE *g, ST *s) {          /* :243-260 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = 0; z < np_ - 1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T bh = P_BH_DLF(z), ch = P_CH_DLF(z);
                size_t p = IZF(z, y, x);
                s->psiHyz_d[p] = (bh * s->psiHyz_d[p]) + (ch * (s->ex[I3(z + 1, y, x)] - s->ex[I3(z, y, x)]) / dx);
                s->psiHxz_d[p] = (bh * s->psiHxz_d[p]) + (ch * (s->ey[I3(z + 1, y, x)] - s->ey[I3(z, y, x)]) / dx);
                s->hy[I3(z, y, x)] -= dtm0 * s->psiHyz_d[p];
                s->hx[I3(z, y, x)] += dtm0 * s->psiHxz_d[p];
            }
}
static void FN(cpml_h_up_)(const GE *g, ST *s) {            /* :263-282 */
    const int ny = g->ny, nx = g->nx, nz = g->nz, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = 0; z < np_ - 1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                int zf = nz - np_ + z;
                T bh = P_BH_URB(z), ch = P_CH_URB(z);
                size_t p = IZF(z, y, x);
                s->psiHyz_u[p] = (bh * s->psiHyz_u[p]) + (ch * (s->ex[I3(zf + 1, y, x)] - s->ex[I3(zf, y, x)]) / dx);
                s->psiHxz_u[p] = (bh * s->psiHxz_u[p]) + (ch * (s->ey[I3(zf + 1, y, x)] - s->ey[I3(zf, y, x)]) / dx);
                s->hy[I3(zf, y, x)] -= dtm0 * s->psiHyz_u[p];
                s->hx[I3(zf, y, x)] += dtm0 * s->psiHxz_u[p];
            }
}
static void FN(cpml_h_left_)(const GE *g, ST *s, int z0, int z1) {   /* :285-302 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < np_ - 1; ++x) {
                T bh = P_BH_DLF(x), ch = P_CH_DLF(x);
                size_t p = IXF(z, y, x);
                s->psiHzx_l[p] = (bh * s->psiHzx_l[p]) + (ch * (s->ey[I3(z, y, x + 1)] - s->ey[I3(z, y, x)]) / dx);
                s->psiHyx_l[p] = (bh * s->psiHyx_l[p]) + (ch * (s->ez[I3(z, y, x + 1)] - s->ez[I3(z, y, x)]) / dx);
                s->hz[I3(z, y, x)] -= dtm0 * s->psiHzx_l[p];
                s->hy[I3(z, y, x)] += dtm0 * s->psiHyx_l[p];
            }
}
static void FN(cpml_h_right_)(const GE *g, ST *s, int z0, int z1) {  /* :305-324 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ny - 1; ++y)
            for (int x = 0; x < np_ - 1; ++x) {
                int xf = nx - np_ + x;
                T bh = P_BH_URB(x), ch = P_CH_URB(x);
                size_t p = IXF(z, y, x);
                s->psiHzx_r[p] = (bh * s->psiHzx_r[p]) + (ch * (s->ey[I3(z, y, xf + 1)] - s->ey[I3(z, y, xf)]) / dx);
                s->psiHyx_r[p] = (bh * s->psiHyx_r[p]) + (ch * (s->ez[I3(z, y, xf + 1)] - s->ez[I3(z, y, xf)]) / dx);
                s->hz[I3(z, y, xf)] -= dtm0 * s->psiHzx_r[p];
                s->hy[I3(z, y, xf)] += dtm0 * s->psiHyx_r[p];
            }
}
static void FN(cpml_h_front_)(const GE *g, ST *s, int z0, int z1) {  /* :327-344 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < np_ - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                T bh = P_BH_DLF(y), ch = P_CH_DLF(y);
                size_t p = IYF(z, y, x);
                s->psiHxy_f[p] = (bh * s->psiHxy_f[p]) + (ch * (s->ez[I3(z, y + 1, x)] - s->ez[I3(z, y, x)]) / dx);
                s->psiHzy_f[p] = (bh * s->psiHzy_f[p]) + (ch * (s->ex[I3(z, y + 1, x)] - s->ex[I3(z, y, x)]) / dx);
                s->hx[I3(z, y, x)] -= dtm0 * s->psiHxy_f[p];
                s->hz[I3(z, y, x)] += dtm0 * s->psiHzy_f[p];
            }
}
static void FN(cpml_h_back_)(const GE *g, ST *s, int z0, int z1) {   /* :347-366 */
    const int ny = g->ny, nx = g->nx, np_ = g->npml; const T dx = g->dx, dtm0 = g->dtm0;
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < np_ - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                int yf = ny - np_ + y;
                T bh = P_BH_URB(y), ch = P_CH_URB(y);
                size_t p = IYF(z, y, x);
                s->psiHxy_b[p] = (bh * s->psiHxy_b[p]) + (ch * (s->ez[I3(z, yf + 1, x)] - s->ez[I3(z, yf, x)]) / dx);
                s->psiHzy_b[p] = (bh * s->psiHzy_b[p]) + (ch * (s->ex[I3(z, yf + 1, x)] - s->ex[I3(z, yf, x)]) / dx);
                s->hx[I3(z, yf, x)] -= dtm0 * s->psiHxy_b[p];
                s->hz[I3(z, yf, x)] += dtm0 * s->psiHzy_b[p];
            }
}

/* ---------------- phasors: cpu/field_updates.py:373-491 -------------------- */
/* acc += w * avg, the running sum read and written in T's complex type, the
 * product and the sum formed in double (numba promotes the 0.5/0.25 literal
 * and the complex exponential to float64/complex128, then stores the sum
 * back to complex64 for T = float).  tw = (cos th, sin th). */
static inline void FN(acc_)(T *acc, double avg, double c, double s_) {
    acc[0] = (T)((double)acc[0] + avg * c);
    acc[1] = (T)((double)acc[1] + avg * s_);
}

/* x plane at px: ey, ez (2-pt), hy, hz (4-pt) -> ph[base..base+3] (nz_ph, ny_ph) */
static void FN(phasors_x_)(const GE *g, ST *s, int px, int base, double ce, double se, double ch, double sh) {
    const int ny = g->ny, nx = g->nx, np_ = g->npml;
    const int nz_ph = g->nz - 2 * np_, ny_ph = ny - 2 * np_;
    for (int z = 1; z < nz_ph; ++z)
        for (int y = 1; y < ny_ph; ++y) {
            size_t q = 2 * ((size_t)z * ny_ph + y);
            int Z = np_ + z, Y = np_ + y;
            T a;
            a = s->ey[I3(Z, Y, px)] + s->ey[I3(Z, Y - 1, px)];
            FN(acc_)(s->ph[base + 0] + q, 0.5 * (double)a, ce, se);
            a = s->ez[I3(Z, Y, px)] + s->ez[I3(Z - 1, Y, px)];
            FN(acc_)(s->ph[base + 1] + q, 0.5 * (double)a, ce, se);
            a = s->hy[I3(Z, Y, px)] + s->hy[I3(Z, Y, px - 1)] + s->hy[I3(Z - 1, Y, px)] + s->hy[I3(Z - 1, Y, px - 1)];
            FN(acc_)(s->ph[base + 2] + q, 0.25 * (double)a, ch, sh);
            a = s->hz[I3(Z, Y, px)] + s->hz[I3(Z, Y, px - 1)] + s->hz[I3(Z, Y - 1, px)] + s->hz[I3(Z, Y - 1, px - 1)];
            FN(acc_)(s->ph[base + 3] + q, 0.25 * (double)a, ch, sh);
        }
}
/* y plane at py: ex, ez (2-pt), hx, hz (4-pt) -> (nz_ph, nx_ph) */
static void FN(phasors_y_)(const GE *g, ST *s, int py, int base, double ce, double se, double ch, double sh) {
    const int ny = g->ny, nx = g->nx, np_ = g->npml;
    const int nz_ph = g->nz - 2 * np_, nx_ph = nx - 2 * np_;
    for (int z = 1; z < nz_ph; ++z)
        for (int x = 1; x < nx_ph; ++x) {
            size_t q = 2 * ((size_t)z * nx_ph + x);
            int Z = np_ + z, X = np_ + x;
            T a;
            a = s->ex[I3(Z, py, X)] + s->ex[I3(Z, py, X - 1)];
            FN(acc_)(s->ph[base + 0] + q, 0.5 * (double)a, ce, se);
            a = s->ez[I3(Z, py, X)] + s->ez[I3(Z - 1, py, X)];
            FN(acc_)(s->ph[base + 1] + q, 0.5 * (double)a, ce, se);
            a = s->hx[I3(Z, py, X)] + s->hx[I3(Z, py - 1, X)] + s->hx[I3(Z - 1, py, X)] + s->hx[I3(Z - 1, py - 1, X)];
            FN(acc_)(s->ph[base + 2] + q, 0.25 * (double)a, ch, sh);
            a = s->hz[I3(Z, py, X)] + s->hz[I3(Z, py - 1, X)] + s->hz[I3(Z, py, X - 1)] + s->hz[I3(Z, py - 1, X - 1)];
            FN(acc_)(s->ph[base + 3] + q, 0.25 * (double)a, ch, sh);
        }
}
/* z plane at pz: ex, ey (2-pt), hx, hy (4-pt) -> (ny_ph, nx_ph).  The
 * incident variant (phasors_incident, :464-491) passes the E and H twiddles
 * swapped. */
static void FN(phasors_z_)(const GE *g, ST *s, int pz, int base, double ce, double se, double ch, double sh) {
    const int ny = g->ny, nx = g->nx, np_ = g->npml;
    const int ny_ph = ny - 2 * np_, nx_ph = nx - 2 * np_;
    for (int y = 1; y < ny_ph; ++y)
        for (int x = 1; x < nx_ph; ++x) {
            size_t q = 2 * ((size_t)y * nx_ph + x);
            int Y = np_ + y, X = np_ + x;
            T a;
            a = s->ex[I3(pz, Y, X)] + s->ex[I3(pz, Y, X - 1)];
            FN(acc_)(s->ph[base + 0] + q, 0.5 * (double)a, ce, se);
            a = s->ey[I3(pz, Y, X)] + s->ey[I3(pz, Y - 1, X)];
            FN(acc_)(s->ph[base + 1] + q, 0.5 * (double)a, ce, se);
            a = s->hx[I3(pz, Y, X)] + s->hx[I3(pz, Y - 1, X)] + s->hx[I3(pz - 1, Y, X)] + s->hx[I3(pz - 1, Y - 1, X)];
            FN(acc_)(s->ph[base + 2] + q, 0.25 * (double)a, ch, sh);
            a = s->hy[I3(pz, Y, X)] + s->hy[I3(pz, Y, X - 1)] + s->hy[I3(pz - 1, Y, X)] + s->hy[I3(pz - 1, Y, X - 1)];
            FN(acc_)(s->ph[base + 3] + q, 0.25 * (double)a, ch, sh);
        }
}

/* ---------------- thread fan-out over z slabs (cpu/fdtd_sim.py:190-208) ---- */
typedef void (*FN(slab_fn_))(const GE *, ST *, int, int);

static void FN(fan_out_)(FN(slab_fn_) f, const GE *g, ST *s, int z0, int z1) {
    int nth = g->nthreads > 0 ? g->nthreads : 1;
    int span = z1 - z0;
    if (nth > span) nth = span;
    int rough = span / nth;
#pragma omp parallel for num_threads(nth) schedule(static, 1)
    for (int i = 0; i < nth; ++i) {
        int a = z0 + i * rough;
        int b = (i == nth - 1) ? z1 : a + rough;
        f(g, s, a, b);
    }
}

static void FN(e_field_update_)(const GE *g, ST *s) {       /* fdtd_sim.py:29-49 */
    FN(fan_out_)(FN(ex_update_), g, s, 1, g->nz - 1);
    FN(fan_out_)(FN(ey_update_), g, s, 1, g->nz - 1);
    FN(fan_out_)(FN(ez_update_), g, s, 0, g->nz - 1);
}
static void FN(h_field_update_)(const GE *g, ST *s) {       /* fdtd_sim.py:82-102 */
    FN(fan_out_)(FN(hx_update_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(hy_update_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(hz_update_), g, s, 1, g->nz - 1);
}
/* CPML faces in the reference order d,u -> l,r -> f,b (fdtd_sim.py:51-80,
 * 104-133).  Paired faces touch disjoint cells, so running them one after
 * the other gives the same values as the reference's concurrent pairs. */
static void FN(e_field_cpml_)(const GE *g, ST *s) {
    FN(cpml_e_down_)(g, s);
    FN(cpml_e_up_)(g, s);
    FN(fan_out_)(FN(cpml_e_left_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_e_right_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_e_front_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_e_back_), g, s, 0, g->nz - 1);
}
static void FN(h_field_cpml_)(const GE *g, ST *s) {
    FN(cpml_h_down_)(g, s);
    FN(cpml_h_up_)(g, s);
    FN(fan_out_)(FN(cpml_h_left_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_h_right_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_h_front_), g, s, 0, g->nz - 1);
    FN(fan_out_)(FN(cpml_h_back_), g, s, 0, g->nz - 1);
}

/* Twiddle of the running DFT: exp(-1j * omega * tt * dt) evaluated as numba
 * does, theta = (-omega * tt) * dt, result (cos theta, sin theta). */
static inline void FN(twiddle_)(double omega, double tt, double dt, double *c, double *s_) {
    double th = (-omega * tt) * dt;
    *c = cos(th);
    *s_ = sin(th);
}

/* One main-simulation time step t (cpu/main_sim.py:140-195). */
void FN(oracle_main_step_)(const GE *g, ST *s, int t) {
    const int nx = g->nx, ny = g->ny, np_ = g->npml;
    const size_t plane = (size_t)ny * nx;
    FN(h_field_update_)(g, s);
    FN(h_field_cpml_)(g, s);
    if (t < g->nt_inc) {                                    /* :147-150 */
        const T *exi = g->inc + (size_t)0 * g->nt_inc * plane + (size_t)t * plane;
        const T *eyi = g->inc + (size_t)1 * g->nt_inc * plane + (size_t)t * plane;
        T *hx = s->hx + (size_t)g->tfsf_z * plane, *hy = s->hy + (size_t)g->tfsf_z * plane;
        for (size_t i = 0; i < plane; ++i) {
            hx[i] += g->Db * eyi[i];
            hy[i] -= g->Db * exi[i];
        }
    }
    FN(e_field_update_)(g, s);
    FN(e_field_cpml_)(g, s);
    if (t < g->nt_inc) {                                    /* :158-161 */
        const T *hxi = g->inc + (size_t)2 * g->nt_inc * plane + (size_t)t * plane;
        const T *hyi = g->inc + (size_t)3 * g->nt_inc * plane + (size_t)t * plane;
        T *ey = s->ey + (size_t)g->tfsf_z * plane, *ex = s->ex + (size_t)g->tfsf_z * plane;
        for (size_t i = 0; i < plane; ++i) {
            ey[i] += g->dte0dx * hxi[i];
            ex[i] -= g->dte0dx * hyi[i];
        }
    }
    double ce, se, ch, sh;
    FN(twiddle_)(g->omega, (double)t, g->dt, &ce, &se);
    FN(twiddle_)(g->omega, (double)t - 0.5, g->dt, &ch, &sh);
    /* :168-193; order of ph[] follows the save list: d, u, l, r, f, b */
    FN(phasors_z_)(g, s, np_ + 1, 0, ce, se, ch, sh);
    FN(phasors_z_)(g, s, g->nz - np_ - 1, 4, ce, se, ch, sh);
    FN(phasors_x_)(g, s, np_ + 1, 8, ce, se, ch, sh);
    FN(phasors_x_)(g, s, nx - np_ - 1, 12, ce, se, ch, sh);
    FN(phasors_y_)(g, s, np_ + 1, 16, ce, se, ch, sh);
    FN(phasors_y_)(g, s, ny - np_ - 1, 20, ce, se, ch, sh);
}

void FN(oracle_main_run_)(const GE *g, ST *s, int t0, int t1) {
    for (int t = t0; t < t1; ++t) FN(oracle_main_step_)(g, s, t);
}

/* Component steps, exposed so tests can compare single passes. */
void FN(oracle_h_half_)(const GE *g, ST *s) { FN(h_field_update_)(g, s); FN(h_field_cpml_)(g, s); }
void FN(oracle_e_half_)(const GE *g, ST *s) { FN(e_field_update_)(g, s); FN(e_field_cpml_)(g, s); }

/* One incident-beam step t (cpu/incident_sim.py:134-161): E, E-CPML, hard
 * source while t < nt_pulse on [1:-1, 0:-1] of the feed-in plane, H, H-CPML,
 * record the read-out plane, incident phasors with E at t-1/2 and H at t.
 * ph[0..3] are the incident phasors (ny_ph, nx_ph). */
void FN(oracle_incident_step_)(const GE *g, ST *s, int t) {
    const int nx = g->nx, ny = g->ny;
    const size_t plane = (size_t)ny * nx;
    FN(e_field_update_)(g, s);
    FN(e_field_cpml_)(g, s);
    if (t < g->nt_inc) {
        const T *sx = g->inc + (size_t)t * plane;
        const T *sy = g->inc + (size_t)g->nt_inc * plane + (size_t)t * plane;
        T *ex = s->ex + (size_t)g->tfsf_z * plane, *ey = s->ey + (size_t)g->tfsf_z * plane;
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 0; x < nx - 1; ++x) {
                ex[(size_t)y * nx + x] = sx[(size_t)y * nx + x];
                ey[(size_t)y * nx + x] = sy[(size_t)y * nx + x];
            }
    }
    FN(h_field_update_)(g, s);
    FN(h_field_cpml_)(g, s);
    double ce, se, ch, sh;
    FN(twiddle_)(g->omega, (double)t - 0.5, g->dt, &ce, &se);
    FN(twiddle_)(g->omega, (double)t, g->dt, &ch, &sh);
    FN(phasors_z_)(g, s, g->read_out, 0, ce, se, ch, sh);
}

/* Incident run: steps [t0, t1); when rec != NULL the read-out planes of
 * ex, ey, hx, hy at step t are written to rec[c][t - t0] with n_rec = t1 - t0. */
void FN(oracle_incident_run_)(const GE *g, ST *s, int t0, int t1) {
    const size_t plane = (size_t)g->ny * g->nx;
    const size_t n_rec = (size_t)(t1 - t0);
    for (int t = t0; t < t1; ++t) {
        FN(oracle_incident_step_)(g, s, t);
        if (g->rec) {
            const T *src[4] = {s->ex, s->ey, s->hx, s->hy};
            for (int c = 0; c < 4; ++c)
                memcpy(g->rec + (size_t)c * n_rec * plane + (size_t)(t - t0) * plane,
                       src[c] + (size_t)g->read_out * plane, plane * sizeof(T));
        }
    }
}

#undef ST
#undef GE
