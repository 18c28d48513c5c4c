/*
 * mhd_oracle.c -- CPU restatement of the reference ideal-MHD finite-volume
 * step of mhd3d3.f90 (tiegcm/mhd3d_mpi_fortran).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product never calls it.
 *
 * Every routine restates one reference routine (cited file:line) and keeps
 * its floating-point operation order (left-to-right sums, divisions kept as
 * divisions, no contraction: build with -ffp-contract=off and no -march), so
 * that on the same inputs it reproduces the flang -fdefault-real-8 build of
 * the reference bit for bit (pinned against the dumps in tests/golden/).
 * The MPI domain decomposition is replaced by one block with the whole
 * domain; the halo exchange plus boundary reflection becomes the same
 * sequential x, y, z ghost fill on that block.
 */
#include "mhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double *d;
    int n1, n2, n3;
} A3;

#define AT(a, i, j, k) \
    ((a).d[((size_t)(i) - 1) + (size_t)(a).n1 * (((size_t)(j) - 1) + (size_t)(a).n2 * ((size_t)(k) - 1))])

static A3 a3_new(int n1, int n2, int n3) {
    A3 a;
    a.d = (double *)calloc((size_t)n1 * n2 * n3, sizeof(double));
    a.n1 = n1;
    a.n2 = n2;
    a.n3 = n3;
    return a;
}
static void a3_free(A3 *a) {
    free(a->d);
    a->d = NULL;
}
static size_t a3_size(const A3 *a) { return (size_t)a->n1 * a->n2 * a->n3; }
static void a3_copy(A3 *dst, const A3 *src) { memcpy(dst->d, src->d, a3_size(src) * sizeof(double)); }

struct orc_state {
    orc_params p;
    int nx, ny, nz;
    double *dx, *dy, *dz; /* dxblock(i,j,k) depends on i only (tensor grid) */
    double tau, tau0;
    A3 rho, vx, vy, vz, pr, bx, by, bz, bi, bj, bk;
    A3 ei, ej, ek;
    A3 flux[ORC_NFLUX];
    A3 rhovx, rhovy, rhovz, eng, rho0, rhovx0, rhovy0, rhovz0, eng0;
    A3 rhop, vxp, vyp, vzp, pp, bxp, byp, bzp, bip, bjp, bkp;
    A3 rhoh, vxh, vyh, vzh, ph, bxh, byh, bzh, bih, bjh, bkh;
    A3 rhovxp, rhovyp, rhovzp, rhovxh, rhovyh, rhovzh, engp;
    /* learned reconstruction in the flux (orc_set_surrogate) */
    float *sur_w;
    double sur_theta;
    int sur_round;
    long sur_faces, sur_fallbacks;
};

static const double PI = 3.141592653589793238462643383279502884; /* = 4*atan(1) in binary64 */

static inline double sq(double x) { return x * x; }
/* SIGN(1.,x) under IEEE semantics: sign bit of x, so -0.0 gives -1. */
static inline double sgn1(double x) { return copysign(1.0, x); }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* --------------------------------------------------------------------- */
/* point stencils (module reconstruct, mhd3d3.f90:827-1599)               */
/* --------------------------------------------------------------------- */

/* interp_pdm 8-point face value, f[0..7] = cells -4..+3 (mhd3d3.f90:871-874) */
static inline double interp8(const double *f) {
    return (-3. * f[0] + 29. * f[1] - 139. * f[2] + 533. * f[3] + 533. * f[4] - 139. * f[5] +
            29. * f[6] - 3. * f[7]) /
           840.;
}

/* PDM2 limiter (mhd3d3.f90:1142-1166) */
static inline void pdm2_pt(double fm2, double fm1, double f0, double fp1, double fi, double pdmb,
                           double *fl, double *fr) {
    double maxf = dmax(fm1, f0), minf = dmin(fm1, f0);
    double fii = dmax(minf, dmin(fi, maxf));
    double dfm = pdmb * (fm1 - fm2), df0 = pdmb * (f0 - fm1), dfp = pdmb * (fp1 - f0);
    double sm = sgn1(dfm), s0 = sgn1(df0), sp = sgn1(dfp);
    dfm = fabs(dfm);
    df0 = fabs(df0);
    dfp = fabs(dfp);
    double qm = fabs(sm + s0), qp = fabs(s0 + sp);
    double dfl = fii - fm1, dfr = f0 - fii;
    *fl = fii - s0 * dmax(0., fabs(dfl) - qm * dfm);
    *fr = fii + s0 * dmax(0., fabs(dfr) - qp * dfp);
}

/* TVD van Leer (mhd3d3.f90:1240-1246) */
static inline void tvd_pt(double fm2, double fm1, double f0, double fp1, double fi, double eps,
                          double *fl, double *fr) {
    double a = (fm1 - fm2) / (f0 - fm1 + eps);
    double b = (a + fabs(a)) / (1. + fabs(a));
    *fl = fm1 - b * (fm1 - fi);
    a = (f0 - fp1) / (fm1 - f0 + eps);
    b = (a + fabs(a)) / (1. + fabs(a));
    *fr = f0 - b * (f0 - fi);
}

/* WENO5 (mhd3d3.f90:1335-1375) */
static inline void weno5_pt(const double *f, double eps, double *fl, double *fr) {
    double fm3 = f[0], fm2 = f[1], fm1 = f[2], f0 = f[3], fp1 = f[4], fp2 = f[5];
    double h0 = (2. * fm3 - 7. * fm2 + 11. * fm1) / 6.;
    double h1 = (-fm2 + 5. * fm1 + 2. * f0) / 6.;
    double h2 = (2. * fm1 + 5. * f0 - fp1) / 6.;
    double s0 = 13. / 12. * sq(fm3 - 2. * fm2 + fm1) + .25 * sq(fm3 - 4. * fm2 + 3. * fm1);
    double s1 = 13. / 12. * sq(fm2 - 2. * fm1 + f0) + .25 * sq(fm2 - f0);
    double s2 = 13. / 12. * sq(fm1 - 2. * f0 + fp1) + .25 * sq(3. * fm1 - 4. * f0 + fp1);
    s0 = .1 / sq(eps + s0);
    s1 = .6 / sq(eps + s1);
    s2 = .3 / sq(eps + s2);
    double t = s0 + s1 + s2;
    double w0 = s0 / t, w1 = s1 / t, w2 = s2 / t;
    *fl = w0 * h0 + w1 * h1 + w2 * h2;

    h0 = (2. * fp2 - 7. * fp1 + 11. * f0) / 6.;
    h1 = (-fp1 + 5. * f0 + 2. * fm1) / 6.;
    h2 = (2. * f0 + 5. * fm1 - fm2) / 6.;
    s0 = 13. / 12. * sq(fp2 - 2. * fp1 + f0) + .25 * sq(fp2 - 4. * fp1 + 3. * f0);
    s1 = 13. / 12. * sq(fp1 - 2. * f0 + fm1) + .25 * sq(fp1 - fm1);
    s2 = 13. / 12. * sq(f0 - 2. * fm1 + fm2) + .25 * sq(3. * f0 - 4. * fm1 + fm2);
    s0 = .1 / sq(eps + s0);
    s1 = .6 / sq(eps + s1);
    s2 = .3 / sq(eps + s2);
    t = s0 + s1 + s2;
    w0 = s0 / t;
    w1 = s1 / t;
    w2 = s2 / t;
    *fr = w0 * h0 + w1 * h1 + w2 * h2;
}

/* minmod / median (mhd3d3.f90:1380-1403) */
static inline double minmod(double a, double b) { return .5 * (sgn1(a) + sgn1(b)) * dmin(fabs(a), fabs(b)); }
static inline double median(double a, double b, double c) { return a + minmod(b - a, c - a); }

/* one side of the MP limiter (mhd3d3.f90:1521-1588); q = (fm3..fp2) mirrored for the right side */
static inline double mp_side(double u, double fm3, double fm2, double fm1, double f0, double fp1,
                             double alpha, double beta) {
    double h0 = fm3 - 2. * fm2 + fm1;
    double h1 = fm2 - 2. * fm1 + f0;
    double h2 = fm1 - 2. * f0 + fp1;
    double s2 = 4. * h1 - h2;
    double s0 = 4. * h2 - h1;
    double s1 = minmod(s2, s0);
    s2 = minmod(h1, h2);
    s0 = minmod(s1, s2);
    double w2 = 4. * h0 - h1;
    double w0 = 4. * h1 - h0;
    double w1 = minmod(w2, w0);
    w2 = minmod(h0, h1);
    w0 = minmod(w1, w2);
    s1 = fm1 + alpha * (fm1 - fm2);
    s2 = .5 * (fm1 + f0 - s0);
    s0 = .5 * (3. * fm1 - fm2) + beta * w0 / 3.;
    w1 = dmin(dmin(fm1, f0), s2);
    w2 = dmin(dmin(fm1, s1), s0);
    w0 = dmax(w1, w2);
    h1 = dmax(dmax(fm1, f0), s2);
    h2 = dmax(dmax(fm1, s1), s0);
    h0 = dmin(h1, h2);
    return median(u, w0, h0);
}

/* MPWENO5 (mhd3d3.f90:1405-1596) */
static inline void mpweno5_pt(const double *f, double eps, double alpha, double beta, double *fl,
                              double *fr) {
    double wl, wr;
    weno5_pt(f, eps, &wl, &wr);
    *fl = mp_side(wl, f[0], f[1], f[2], f[3], f[4], alpha, beta);
    *fr = mp_side(wr, f[5], f[4], f[3], f[2], f[1], alpha, beta);
}

/* get_v_face (mhd3d3.f90:2594-2615) */
static inline void vface_pt(double lam, double v, double *v0p, double *v0n, double *v1p, double *v1n) {
    double a = sqrt(lam);
    double b = .5 * exp(-lam * (v * v)) / a / sqrt(PI);
    *v0p = .5 * erfc(-a * v);
    *v0n = .5 * erfc(a * v);
    *v1p = v * *v0p + b;
    *v1n = v * *v0n - b;
}

/* --------------------------------------------------------------------- */
/* stencil gather shared by all reconstructions                           */
/* --------------------------------------------------------------------- */

/* f at stencil offset m of output point (i,j,k) of reconstruct direction
 * dir (1-3 truncated with offset ng in all axes, 4-6 untruncated in the two
 * transverse axes), with the PDM2 shifts (sx,sy,sz)  (mhd3d3.f90:1089-1138) */
static inline double gat(const A3 *f, int ng, int dir, int i, int j, int k, int m, int sx, int sy, int sz) {
    switch (dir) {
    case 1: return AT(*f, ng + i + m + sx, ng + j + sy, ng + k + sz);
    case 2: return AT(*f, ng + i + sx, ng + j + m + sy, ng + k + sz);
    case 3: return AT(*f, ng + i + sx, ng + j + sy, ng + k + m + sz);
    case 4: return AT(*f, ng + i + m + sx, j, k);
    case 5: return AT(*f, i, ng + j + m + sy, k);
    default: return AT(*f, i, j, ng + k + m + sz);
    }
}

/* reconstruct_3d at one point (mhd3d3.f90:833-860) */
static void recon_pt(const orc_params *p, const A3 *f, int dir, int i, int j, int k, double *fl, double *fr) {
    int ng = p->ng;
    double s[8];
    switch (p->reconop) {
    case 1: {
        for (int m = 0; m < 8; ++m) s[m] = gat(f, ng, dir, i, j, k, m - 4, 0, 0, 0);
        double fi = interp8(s);
        pdm2_pt(s[2], s[3], s[4], s[5], fi, p->pdmb, fl, fr);
        break;
    }
    case 2: {
        for (int m = 0; m < 4; ++m) s[m] = gat(f, ng, dir, i, j, k, m - 2, 0, 0, 0);
        double fi = (s[1] + s[2]) / 2.;
        tvd_pt(s[0], s[1], s[2], s[3], fi, p->eps, fl, fr);
        break;
    }
    case 3:
        for (int m = 0; m < 6; ++m) s[m] = gat(f, ng, dir, i, j, k, m - 3, 0, 0, 0);
        weno5_pt(s, p->eps, fl, fr);
        break;
    case 4:
        for (int m = 0; m < 6; ++m) s[m] = gat(f, ng, dir, i, j, k, m - 3, 0, 0, 0);
        mpweno5_pt(s, p->eps, p->alpha, p->beta, fl, fr);
        break;
    default: *fl = *fr = 0.; break;
    }
}

void orc_reconstruct(const orc_params *p, const double *f, int fn1, int fn2, int fn3, int lnx, int lny,
                     int lnz, int dir, double *fl, double *fr) {
    A3 fa = {(double *)f, fn1, fn2, fn3};
    A3 la = {fl, lnx, lny, lnz}, ra = {fr, lnx, lny, lnz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= lnz; ++k)
        for (int j = 1; j <= lny; ++j)
            for (int i = 1; i <= lnx; ++i) recon_pt(p, &fa, dir, i, j, k, &AT(la, i, j, k), &AT(ra, i, j, k));
}

void orc_v_face(int n, const double *lam, const double *v, double *v0p, double *v0n, double *v1p, double *v1n) {
    for (int q = 0; q < n; ++q) vface_pt(lam[q], v[q], &v0p[q], &v0n[q], &v1p[q], &v1n[q]);
}

/* --------------------------------------------------------------------- */
/* corner velocity  get_vav_center2corner (mhd3d3.f90:1858-2116)          */
/* --------------------------------------------------------------------- */

/* lv at base cell (I,J,K) + o1 along axis a1 + o2 along axis a2 (axes 1..3) */
static inline double cv(const A3 *lv, int I, int J, int K, int a1, int o1, int a2, int o2) {
    int c[4] = {0, I, J, K};
    c[a1] += o1;
    c[a2] += o2;
    return AT(*lv, c[1], c[2], c[3]);
}

/* one 1-D reconstruction of lv along axis a, at transverse offset ot along
 * axis b, of the face between offsets -1 and 0 along a (cell-centred
 * stencils; used by the PDM and TVD corner branches) */
static double vav_pt(const orc_params *p, const A3 *lv, int dir, int i, int j, int k) {
    int ng = p->ng;
    int d1, d2;
    if (dir == 1) { d1 = 2; d2 = 3; }
    else if (dir == 2) { d1 = 3; d2 = 1; }
    else { d1 = 1; d2 = 2; }
    const int I = ng + i, J = ng + j, K = ng + k;
    double v00_0 = 0., v10_0 = 0., v01_0 = 0., v11_0 = 0., v00_1 = 0., v01_1 = 0., v10_1 = 0., v11_1 = 0.;
    double s[8];

    if (p->reconop == 1) {
        /* interp2_pdm: first along d1 at 8 levels of d2, then along d2 (926-989) */
        double g[8];
        for (int m2 = 0; m2 < 8; ++m2) {
            for (int m1 = 0; m1 < 8; ++m1) s[m1] = cv(lv, I, J, K, d1, m1 - 4, d2, m2 - 4);
            g[m2] = interp8(s);
        }
        double vi = interp8(g);
        /* PDM2 calls (1899-1912) */
        pdm2_pt(cv(lv, I, J, K, d1, -2, d2, -1), cv(lv, I, J, K, d1, -1, d2, -1), cv(lv, I, J, K, d1, 0, d2, -1),
                cv(lv, I, J, K, d1, 1, d2, -1), vi, p->pdmb, &v00_0, &v10_0);
        pdm2_pt(cv(lv, I, J, K, d2, -2, d1, -1), cv(lv, I, J, K, d2, -1, d1, -1), cv(lv, I, J, K, d2, 0, d1, -1),
                cv(lv, I, J, K, d2, 1, d1, -1), vi, p->pdmb, &v00_1, &v01_1);
        pdm2_pt(cv(lv, I, J, K, d2, -2, d1, 0), cv(lv, I, J, K, d2, -1, d1, 0), cv(lv, I, J, K, d2, 0, d1, 0),
                cv(lv, I, J, K, d2, 1, d1, 0), vi, p->pdmb, &v10_1, &v11_1);
        pdm2_pt(cv(lv, I, J, K, d1, -2, d2, 0), cv(lv, I, J, K, d1, -1, d2, 0), cv(lv, I, J, K, d1, 0, d2, 0),
                cv(lv, I, J, K, d1, 1, d2, 0), vi, p->pdmb, &v01_0, &v11_0);
    } else if (p->reconop == 2) {
        /* interp2_tvd (1037-1067) */
        double vi = (cv(lv, I, J, K, d1, -1, d2, -1) + cv(lv, I, J, K, d1, 0, d2, -1) + cv(lv, I, J, K, d1, 0, d2, 0) +
                     cv(lv, I, J, K, d1, -1, d2, 0)) /
                    4.;
        /* TVD ignores its shift arguments (1187-1238): every sweep sits at
         * transverse offset 0 (1915-1933) */
        tvd_pt(cv(lv, I, J, K, d1, -2, d2, 0), cv(lv, I, J, K, d1, -1, d2, 0), cv(lv, I, J, K, d1, 0, d2, 0),
               cv(lv, I, J, K, d1, 1, d2, 0), vi, p->eps, &v00_0, &v10_0);
        tvd_pt(cv(lv, I, J, K, d2, -2, d1, 0), cv(lv, I, J, K, d2, -1, d1, 0), cv(lv, I, J, K, d2, 0, d1, 0),
               cv(lv, I, J, K, d2, 1, d1, 0), vi, p->eps, &v00_1, &v01_1);
        v10_1 = v00_1;
        v11_1 = v01_1;
        v01_0 = v00_0;
        v11_0 = v10_0;
    } else {
        /* WENO5 / MPWENO5 two-sweep corner values (1936-2051).
         * family 1: sweep along d1 first, then d2; family 0: d2 first, then
         * d1.  The first sweep's left/right state becomes the first label
         * digit and the second sweep's the second digit in both families. */
        double L[6], R[6], t[6];
        for (int fam = 1; fam >= 0; --fam) {
            int a = fam ? d1 : d2, b = fam ? d2 : d1;
            for (int mb = 0; mb < 6; ++mb) {
                for (int ma = 0; ma < 6; ++ma) t[ma] = cv(lv, I, J, K, a, ma - 3, b, mb - 3);
                if (p->reconop == 3) weno5_pt(t, p->eps, &L[mb], &R[mb]);
                else mpweno5_pt(t, p->eps, p->alpha, p->beta, &L[mb], &R[mb]);
            }
            double xl, xr, yl, yr;
            if (p->reconop == 3) {
                weno5_pt(L, p->eps, &xl, &xr);
                weno5_pt(R, p->eps, &yl, &yr);
            } else {
                mpweno5_pt(L, p->eps, p->alpha, p->beta, &xl, &xr);
                mpweno5_pt(R, p->eps, p->alpha, p->beta, &yl, &yr);
            }
            if (fam) { v00_1 = xl; v01_1 = xr; v10_1 = yl; v11_1 = yr; }
            else { v00_0 = xl; v01_0 = xr; v10_0 = yl; v11_0 = yr; }
        }
    }

    /* selection by the larger one-sided jump (2058-2106) */
    double dif00 = fabs(cv(lv, I, J, K, d1, -2, d2, -1) - cv(lv, I, J, K, d1, -1, d2, -1)) -
                   fabs(cv(lv, I, J, K, d1, -1, d2, -2) - cv(lv, I, J, K, d1, -1, d2, -1));
    double dif10 = fabs(cv(lv, I, J, K, d1, 1, d2, -1) - cv(lv, I, J, K, d1, 0, d2, -1)) -
                   fabs(cv(lv, I, J, K, d1, 0, d2, -2) - cv(lv, I, J, K, d1, 0, d2, -1));
    double dif01 = fabs(cv(lv, I, J, K, d1, -2, d2, 0) - cv(lv, I, J, K, d1, -1, d2, 0)) -
                   fabs(cv(lv, I, J, K, d1, -1, d2, 1) - cv(lv, I, J, K, d1, -1, d2, 0));
    double dif11 = fabs(cv(lv, I, J, K, d1, 1, d2, 0) - cv(lv, I, J, K, d1, 0, d2, 0)) -
                   fabs(cv(lv, I, J, K, d1, 0, d2, 1) - cv(lv, I, J, K, d1, 0, d2, 0));
    dif00 = sgn1(dif00);
    dif10 = sgn1(dif10);
    dif01 = sgn1(dif01);
    dif11 = sgn1(dif11);
    double v00 = (1. + dif00) / 2. * v00_0 + (1. - dif00) / 2. * v00_1;
    double v10 = (1. + dif10) / 2. * v10_0 + (1. - dif10) / 2. * v10_1;
    double v01 = (1. + dif01) / 2. * v01_0 + (1. - dif01) / 2. * v01_1;
    double v11 = (1. + dif11) / 2. * v11_0 + (1. - dif11) / 2. * v11_1;
    return .25 * (v00 + v01 + v10 + v11);
}

/* --------------------------------------------------------------------- */
/* get_eijk_LFM (mhd3d3.f90:1754-1856)                                    */
/* --------------------------------------------------------------------- */
static void eijk_lfm(orc_state *s, const A3 *lrho, const A3 *lvx, const A3 *lvy, const A3 *lvz, const A3 *lbx,
                     const A3 *lby, const A3 *lbz, const A3 *lbi, const A3 *lbj, const A3 *lbk) {
    const orc_params *p = &s->p;
    const int ng = p->ng;
    A3 vf = a3_new(s->nx, s->ny, s->nz);
    size_t n = a3_size(&vf);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n; ++q)
        vf.d[q] = sqrt(sq(lvx->d[q]) + sq(lvy->d[q]) + sq(lvz->d[q])) +
                  sqrt((sq(lbx->d[q]) + sq(lby->d[q]) + sq(lbz->d[q])) / lrho->d[q]);

    for (int kkk = 1; kkk <= 3; ++kkk) {
        int ii = p->nx0 + 1, jj = p->ny0 + 1, kk = p->nz0 + 1, d1, d2;
        const A3 *v1, *v2, *b1, *b2;
        A3 *out;
        if (kkk == 1) { ii = p->nx0; v1 = lvy; v2 = lvz; b1 = lbj; b2 = lbk; d1 = 3; d2 = 2; out = &s->ei; }
        else if (kkk == 2) { jj = p->ny0; v1 = lvz; v2 = lvx; b1 = lbk; b2 = lbi; d1 = 1; d2 = 3; out = &s->ej; }
        else { kk = p->nz0; v1 = lvx; v2 = lvy; b1 = lbi; b2 = lbj; d1 = 2; d2 = 1; out = &s->ek; }
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
        for (int k = 1; k <= kk; ++k)
            for (int j = 1; j <= jj; ++j)
                for (int i = 1; i <= ii; ++i) {
                    double v1av = vav_pt(p, v1, kkk, i, j, k);
                    double v2av = vav_pt(p, v2, kkk, i, j, k);
                    double sig1 = sgn1(v1av), sig2 = sgn1(v2av);
                    double b1l, b1r, b2l, b2r;
                    recon_pt(p, b1, d1, i, j, k, &b1l, &b1r);
                    recon_pt(p, b2, d2, i, j, k, &b2l, &b2r);
                    double b1u = (1. + sig2) / 2. * b1l + (1. - sig2) / 2. * b1r;
                    double b2u = (1. + sig1) / 2. * b2l + (1. - sig1) / 2. * b2r;
                    double dum = -v1av * b2u + v2av * b1u;
                    double eta;
                    if (kkk == 1)
                        eta = .25 * (AT(vf, i + ng, j + ng - 1, k + ng - 1) + AT(vf, i + ng, j + ng, k + ng - 1) +
                                     AT(vf, i + ng, j + ng, k + ng) + AT(vf, i + ng, j + ng - 1, k + ng));
                    else if (kkk == 2)
                        eta = .25 * (AT(vf, i + ng - 1, j + ng, k + ng - 1) + AT(vf, i + ng - 1, j + ng, k + ng) +
                                     AT(vf, i + ng, j + ng, k + ng) + AT(vf, i + ng, j + ng, k + ng - 1));
                    else
                        eta = .25 * (AT(vf, i + ng - 1, j + ng - 1, k + ng) + AT(vf, i + ng, j + ng - 1, k + ng) +
                                     AT(vf, i + ng, j + ng, k + ng) + AT(vf, i + ng - 1, j + ng, k + ng));
                    dum = dum + eta * (b1l + b2r - b1r - b2l);
                    AT(*out, i, j, k) = dum;
                }
    }
    a3_free(&vf);
}

/* --------------------------------------------------------------------- */
/* get_flux_stress (mhd3d3.f90:2385-2560)                                 */
/* --------------------------------------------------------------------- */
static int sur_face(const orc_state *s, const A3 *const *f, int dir, int i, int j, int k, double *lv,
                   double *rv);
static void flux_stress(orc_state *s, const A3 *lrho, const A3 *lvx, const A3 *lvy, const A3 *lvz, const A3 *lp,
                        const A3 *lbx, const A3 *lby, const A3 *lbz) {
    const orc_params *p = &s->p;
    const double gam = p->gam;
    for (int kkk = 1; kkk <= 3; ++kkk) {
        int ii = p->nx0 + (kkk == 1), jj = p->ny0 + (kkk == 2), kk = p->nz0 + (kkk == 3);
        A3 *F = &s->flux[8 * (kkk - 1)];
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= kk; ++k)
            for (int j = 1; j <= jj; ++j)
                for (int i = 1; i <= ii; ++i) {
                    double rl, rr, vxl, vxr, vyl, vyr, vzl, vzr, pl, prr, bxl, bxr, byl, byr, bzl, bzr;
                    int net = 0;
                    if (s->sur_w) {
                        const A3 *fv[8] = {lrho, lvx, lvy, lvz, lp, lbx, lby, lbz};
                        double lv[8], rv[8];
                        net = sur_face(s, fv, kkk, i, j, k, lv, rv);
                        rl = lv[0], rr = rv[0], vxl = lv[1], vxr = rv[1], vyl = lv[2], vyr = rv[2];
                        vzl = lv[3], vzr = rv[3], pl = lv[4], prr = rv[4], bxl = lv[5], bxr = rv[5];
                        byl = lv[6], byr = rv[6], bzl = lv[7], bzr = rv[7];
                    }
                    if (!net) {
                        recon_pt(p, lrho, kkk, i, j, k, &rl, &rr);
                        recon_pt(p, lvx, kkk, i, j, k, &vxl, &vxr);
                        recon_pt(p, lvy, kkk, i, j, k, &vyl, &vyr);
                        recon_pt(p, lvz, kkk, i, j, k, &vzl, &vzr);
                        recon_pt(p, lp, kkk, i, j, k, &pl, &prr);
                        recon_pt(p, lbx, kkk, i, j, k, &bxl, &bxr);
                        recon_pt(p, lby, kkk, i, j, k, &byl, &byr);
                        recon_pt(p, lbz, kkk, i, j, k, &bzl, &bzr);
                    }
                    double v0p, v0n, v1p, v1n;
                    /* left */
                    double leng = .5 * rl * (sq(vxl) + sq(vyl) + sq(vzl)) + pl / (gam - 1.);
                    double llam = rl / (2. * pl);
                    double vnl = kkk == 1 ? vxl : (kkk == 2 ? vyl : vzl);
                    vface_pt(llam, vnl, &v0p, &v0n, &v1p, &v1n);
                    double rp = rl * v1p, rvxp = rl * vxl * v1p, rvyp = rl * vyl * v1p, rvzp = rl * vzl * v1p;
                    double ep = (leng + .5 * pl) * v1p;
                    if (kkk == 1) rvxp = rvxp + pl * v0p;
                    else if (kkk == 2) rvyp = rvyp + pl * v0p;
                    else rvzp = rvzp + pl * v0p;
                    ep = ep + .5 * pl * vnl * v0p;
                    double lbsq = sq(bxl) + sq(byl) + sq(bzl);
                    llam = rl / (2. * pl + lbsq);
                    vface_pt(llam, vnl, &v0p, &v0n, &v1p, &v1n);
                    double bnl = kkk == 1 ? bxl : (kkk == 2 ? byl : bzl);
                    double sxp = (kkk == 1 ? (.5 * lbsq - bxl * bxl) : (-bxl * bnl)) * v0p;
                    double syp = (kkk == 2 ? (.5 * lbsq - byl * byl) : (-byl * bnl)) * v0p;
                    double szp = (kkk == 3 ? (.5 * lbsq - bzl * bzl) : (-bzl * bnl)) * v0p;
                    /* right */
                    leng = .5 * rr * (sq(vxr) + sq(vyr) + sq(vzr)) + prr / (gam - 1.);
                    llam = rr / (2. * prr);
                    double vnr = kkk == 1 ? vxr : (kkk == 2 ? vyr : vzr);
                    vface_pt(llam, vnr, &v0p, &v0n, &v1p, &v1n);
                    double rn = rr * v1n, rvxn = rr * vxr * v1n, rvyn = rr * vyr * v1n, rvzn = rr * vzr * v1n;
                    double en = (leng + .5 * prr) * v1n;
                    if (kkk == 1) rvxn = rvxn + prr * v0n;
                    else if (kkk == 2) rvyn = rvyn + prr * v0n;
                    else rvzn = rvzn + prr * v0n;
                    en = en + .5 * prr * vnr * v0n;
                    lbsq = sq(bxr) + sq(byr) + sq(bzr);
                    llam = rr / (2. * prr + lbsq);
                    vface_pt(llam, vnr, &v0p, &v0n, &v1p, &v1n);
                    double bnr = kkk == 1 ? bxr : (kkk == 2 ? byr : bzr);
                    double sxn = (kkk == 1 ? (.5 * lbsq - bxr * bxr) : (-bxr * bnr)) * v0n;
                    double syn = (kkk == 2 ? (.5 * lbsq - byr * byr) : (-byr * bnr)) * v0n;
                    double szn = (kkk == 3 ? (.5 * lbsq - bzr * bzr) : (-bzr * bnr)) * v0n;
                    AT(F[0], i, j, k) = rp + rn;
                    AT(F[1], i, j, k) = rvxp + rvxn;
                    AT(F[2], i, j, k) = rvyp + rvyn;
                    AT(F[3], i, j, k) = rvzp + rvzn;
                    AT(F[4], i, j, k) = ep + en;
                    AT(F[5], i, j, k) = sxp + sxn;
                    AT(F[6], i, j, k) = syp + syn;
                    AT(F[7], i, j, k) = szp + szn;
                }
    }
}

/* --------------------------------------------------------------------- */
/* ghost fill: module mpi_fillers + boundary (mhd3d3.f90:346-616,781-819) */
/* --------------------------------------------------------------------- */

/* fill the ng ghost layers of array a along `axis` (1..3).  face != 0 when
 * the array is face-centred along that axis.  n0 = interior cells. */
static void fill_axis(A3 *a, int axis, int face, int n0, int ng, int bc) {
    int n[4] = {0, a->n1, a->n2, a->n3};
    int L = n[axis];
    int nord = 2 * ng;
    int o1 = axis == 1 ? 2 : 1, o2 = axis == 3 ? 2 : 3;
    for (int c2 = 1; c2 <= n[o2]; ++c2)
        for (int c1 = 1; c1 <= n[o1]; ++c1)
            for (int g = 1; g <= ng; ++g) {
                int dstL = g, srcL, dstR, srcR;
                if (!face) {
                    dstR = L + 1 - g;
                    if (bc == 2) { srcL = L - nord + g; dstR = L - ng + g; srcR = ng + g; }
                    else { srcL = nord + 1 - g; srcR = L - nord + g; }
                } else {
                    if (bc == 2) { srcL = n0 + g; dstR = L - ng + g; srcR = ng + 1 + g; }
                    else { srcL = nord + 2 - g; dstR = L + 1 - g; srcR = n0 + g; }
                }
                int c[4];
                c[o1] = c1; c[o2] = c2;
                c[axis] = srcL; double vL = AT(*a, c[1], c[2], c[3]);
                c[axis] = srcR; double vR = AT(*a, c[1], c[2], c[3]);
                c[axis] = dstL; AT(*a, c[1], c[2], c[3]) = vL;
                c[axis] = dstR; AT(*a, c[1], c[2], c[3]) = vR;
            }
}

void orc_fill(orc_state *s) {
    const orc_params *p = &s->p;
    int n0[4] = {0, p->nx0, p->ny0, p->nz0};
    A3 *hyd[5] = {&s->rho, &s->pr, &s->vx, &s->vy, &s->vz};
    for (int ax = 1; ax <= 3; ++ax)
        for (int q = 0; q < 5; ++q) fill_axis(hyd[q], ax, 0, n0[ax], p->ng, p->bc[ax - 1]);
    A3 *mag[3] = {&s->bi, &s->bj, &s->bk};
    for (int ax = 1; ax <= 3; ++ax)
        for (int q = 0; q < 3; ++q) fill_axis(mag[q], ax, q + 1 == ax, n0[ax], p->ng, p->bc[ax - 1]);
}

/* get_bxyz (mhd3d3.f90:2118-2128) */
void orc_bxyz(orc_state *s) {
    int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(s->bx, i, j, k) = .5 * (AT(s->bi, i + 1, j, k) + AT(s->bi, i, j, k));
                AT(s->by, i, j, k) = .5 * (AT(s->bj, i, j + 1, k) + AT(s->bj, i, j, k));
                AT(s->bz, i, j, k) = .5 * (AT(s->bk, i, j, k + 1) + AT(s->bk, i, j, k));
            }
}

/* get_dt (mhd3d3.f90:1608-1639): min over the ghost-inclusive block */
double orc_get_dt(orc_state *s) {
    const double gam = s->p.gam, cfl = s->p.cfl;
    double tmin = INFINITY;
#pragma omp parallel for reduction(min : tmin) schedule(static)
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                double v = sqrt(sq(AT(s->vx, i, j, k)) + sq(AT(s->vy, i, j, k)) + sq(AT(s->vz, i, j, k)));
                double btot = sqrt(sq(AT(s->bx, i, j, k)) + sq(AT(s->by, i, j, k)) + sq(AT(s->bz, i, j, k)));
                double valf = btot / sqrt(AT(s->rho, i, j, k));
                double vs = sqrt(gam * AT(s->pr, i, j, k) / AT(s->rho, i, j, k));
                double vcfl = v + sqrt(sq(valf) + sq(vs));
                double dt = cfl / (vcfl / s->dx[i - 1] + vcfl / s->dy[j - 1] + vcfl / s->dz[k - 1]);
                if (dt < tmin) tmin = dt;
            }
    s->tau = tmin;
    return tmin;
}

/* get_conserve (mhd3d3.f90:1641-1650) */
static void get_conserve(orc_state *s) {
    const double gam = s->p.gam;
    size_t n = a3_size(&s->rho);
    for (size_t q = 0; q < n; ++q) {
        double r = s->rho.d[q], u = s->vx.d[q], v = s->vy.d[q], w = s->vz.d[q];
        s->rho0.d[q] = r;
        s->rhovx0.d[q] = r * u;
        s->rhovy0.d[q] = r * v;
        s->rhovz0.d[q] = r * w;
        s->eng0.d[q] = .5 * r * (sq(u) + sq(v) + sq(w)) + s->pr.d[q] / (gam - 1.);
    }
}

/* advance_corrector1 (mhd3d3.f90:3842-3878) */
static void corrector1(orc_state *s) {
    const orc_params *p = &s->p;
    const int ng = p->ng;
    const double tau = s->tau;
    A3 *F = s->flux;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= p->nz0; ++k)
        for (int j = 1; j <= p->ny0; ++j)
            for (int i = 1; i <= p->nx0; ++i) {
                double dxi = s->dx[ng + i - 1], dyj = s->dy[ng + j - 1], dzk = s->dz[ng + k - 1];
                A3 *U0[5] = {&s->rho0, &s->eng0, &s->rhovx0, &s->rhovy0, &s->rhovz0};
                A3 *U[5] = {&s->rho, &s->eng, &s->rhovx, &s->rhovy, &s->rhovz};
                const int fq[5] = {0, 4, 1, 2, 3};
                for (int q = 0; q < 5; ++q) {
                    int f = fq[q];
                    AT(*U[q], ng + i, ng + j, ng + k) =
                        AT(*U0[q], ng + i, ng + j, ng + k) -
                        tau * ((AT(F[f], i + 1, j, k) - AT(F[f], i, j, k)) / dxi +
                               (AT(F[8 + f], i, j + 1, k) - AT(F[8 + f], i, j, k)) / dyj +
                               (AT(F[16 + f], i, j, k + 1) - AT(F[16 + f], i, j, k)) / dzk);
                }
            }
}

/* advance_corrector2 (mhd3d3.f90:3880-3907) */
static void corrector2(orc_state *s) {
    const orc_params *p = &s->p;
    const int ng = p->ng;
    const double tau = s->tau;
    A3 *F = s->flux;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= p->nz0; ++k)
        for (int j = 1; j <= p->ny0; ++j)
            for (int i = 1; i <= p->nx0; ++i) {
                double dxi = s->dx[ng + i - 1], dyj = s->dy[ng + j - 1], dzk = s->dz[ng + k - 1];
                A3 *U[3] = {&s->rhovx, &s->rhovy, &s->rhovz};
                for (int q = 0; q < 3; ++q) {
                    int f = 5 + q;
                    AT(*U[q], ng + i, ng + j, ng + k) =
                        AT(*U[q], ng + i, ng + j, ng + k) -
                        tau * ((AT(F[f], i + 1, j, k) - AT(F[f], i, j, k)) / dxi +
                               (AT(F[8 + f], i, j + 1, k) - AT(F[8 + f], i, j, k)) / dyj +
                               (AT(F[16 + f], i, j, k + 1) - AT(F[16 + f], i, j, k)) / dzk);
                }
            }
}

/* advance_bijk (mhd3d3.f90:3909-3933) */
static void advance_bijk(orc_state *s) {
    const orc_params *p = &s->p;
    const int ng = p->ng;
    const double tau = s->tau;
    for (int k = 1; k <= p->nz0; ++k)
        for (int j = 1; j <= p->ny0; ++j)
            for (int i = 1; i <= p->nx0 + 1; ++i)
                AT(s->bi, ng + i, ng + j, ng + k) =
                    AT(s->bi, ng + i, ng + j, ng + k) -
                    tau * ((AT(s->ek, i, j + 1, k) - AT(s->ek, i, j, k)) / s->dy[ng + j - 1] -
                           (AT(s->ej, i, j, k + 1) - AT(s->ej, i, j, k)) / s->dz[ng + k - 1]);
    for (int k = 1; k <= p->nz0; ++k)
        for (int j = 1; j <= p->ny0 + 1; ++j)
            for (int i = 1; i <= p->nx0; ++i)
                AT(s->bj, ng + i, ng + j, ng + k) =
                    AT(s->bj, ng + i, ng + j, ng + k) -
                    tau * ((AT(s->ei, i, j, k + 1) - AT(s->ei, i, j, k)) / s->dz[ng + k - 1] -
                           (AT(s->ek, i + 1, j, k) - AT(s->ek, i, j, k)) / s->dx[ng + i - 1]);
    for (int k = 1; k <= p->nz0 + 1; ++k)
        for (int j = 1; j <= p->ny0; ++j)
            for (int i = 1; i <= p->nx0; ++i)
                AT(s->bk, ng + i, ng + j, ng + k) =
                    AT(s->bk, ng + i, ng + j, ng + k) -
                    tau * ((AT(s->ej, i + 1, j, k) - AT(s->ej, i, j, k)) / s->dx[ng + i - 1] -
                           (AT(s->ei, i, j + 1, k) - AT(s->ei, i, j, k)) / s->dy[ng + j - 1]);
}

/* advance_previous (mhd3d3.f90:3794-3819) */
static void advance_previous(orc_state *s) {
    const double gam = s->p.gam;
    a3_copy(&s->rhop, &s->rho);
    a3_copy(&s->bip, &s->bi);
    a3_copy(&s->bjp, &s->bj);
    a3_copy(&s->bkp, &s->bk);
    if (s->p.stepop == 1) {
        a3_copy(&s->vxp, &s->vx);
        a3_copy(&s->vyp, &s->vy);
        a3_copy(&s->vzp, &s->vz);
        a3_copy(&s->pp, &s->pr);
        a3_copy(&s->bxp, &s->bx);
        a3_copy(&s->byp, &s->by);
        a3_copy(&s->bzp, &s->bz);
    } else {
        size_t n = a3_size(&s->rho);
        for (size_t q = 0; q < n; ++q) {
            double r = s->rho.d[q], u = s->vx.d[q], v = s->vy.d[q], w = s->vz.d[q];
            s->rhovxp.d[q] = r * u;
            s->rhovyp.d[q] = r * v;
            s->rhovzp.d[q] = r * w;
            s->engp.d[q] = .5 * r * (sq(u) + sq(v) + sq(w)) + s->pr.d[q] / (gam - 1.);
        }
    }
}

/* interior-cell loop helper */
#define FOR_INTERIOR(s, ...)                                                   \
    for (int k_ = (s)->p.ng + 1; k_ <= (s)->p.ng + (s)->p.nz0; ++k_)          \
        for (int j_ = (s)->p.ng + 1; j_ <= (s)->p.ng + (s)->p.ny0; ++j_)      \
            for (int i_ = (s)->p.ng + 1; i_ <= (s)->p.ng + (s)->p.nx0; ++i_) { \
                size_t q = (size_t)(i_ - 1) + (size_t)(s)->nx * ((size_t)(j_ - 1) + (size_t)(s)->ny * (k_ - 1)); \
                __VA_ARGS__                                                     \
            }

/* pressure from the hydro-only momentum (mhd3d3.f90:3685, 3876) */
static inline double prs(double gam, double eng, double mx, double my, double mz, double r) {
    return (gam - 1.) * (eng - .5 * (sq(mx) + sq(my) + sq(mz)) / r);
}

/* one TVD RK stage's update after the flux (mhd3d3.f90:3669-3707 for RK2,
 * 3737-3785 for RK3): corrector1, stage combine and pressure, corrector2,
 * combine and velocity, advance_bijk and the face combine */
static void rk_update(orc_state *s, int nst, int st) {
    const double gam = s->p.gam;
    {
        corrector1(s);
        FOR_INTERIOR(s, {
            if (st == 1) {
                s->rhovxh.d[q] = s->rhovx.d[q];
                s->rhovyh.d[q] = s->rhovy.d[q];
                s->rhovzh.d[q] = s->rhovz.d[q];
            } else if (nst == 2 || st == 2) {
                if (nst == 2) {
                    s->rho.d[q] = (s->rhop.d[q] + s->rho.d[q]) / 2.;
                    s->rhovxh.d[q] = (s->rhovxp.d[q] + s->rhovx.d[q]) / 2.;
                    s->rhovyh.d[q] = (s->rhovyp.d[q] + s->rhovy.d[q]) / 2.;
                    s->rhovzh.d[q] = (s->rhovzp.d[q] + s->rhovz.d[q]) / 2.;
                    s->eng.d[q] = (s->engp.d[q] + s->eng.d[q]) / 2.;
                } else {
                    s->rho.d[q] = (3. * s->rhop.d[q] + s->rho.d[q]) / 4.;
                    s->rhovxh.d[q] = (3. * s->rhovxp.d[q] + s->rhovx.d[q]) / 4.;
                    s->rhovyh.d[q] = (3. * s->rhovyp.d[q] + s->rhovy.d[q]) / 4.;
                    s->rhovzh.d[q] = (3. * s->rhovzp.d[q] + s->rhovz.d[q]) / 4.;
                    s->eng.d[q] = (3. * s->engp.d[q] + s->eng.d[q]) / 4.;
                }
            } else {
                s->rho.d[q] = (s->rhop.d[q] + 2. * s->rho.d[q]) / 3.;
                s->rhovxh.d[q] = (s->rhovxp.d[q] + 2. * s->rhovx.d[q]) / 3.;
                s->rhovyh.d[q] = (s->rhovyp.d[q] + 2. * s->rhovy.d[q]) / 3.;
                s->rhovzh.d[q] = (s->rhovzp.d[q] + 2. * s->rhovz.d[q]) / 3.;
                s->eng.d[q] = (s->engp.d[q] + 2. * s->eng.d[q]) / 3.;
            }
            s->pr.d[q] = prs(gam, s->eng.d[q], s->rhovxh.d[q], s->rhovyh.d[q], s->rhovzh.d[q], s->rho.d[q]);
        });
        corrector2(s);
        FOR_INTERIOR(s, {
            A3 *m[3] = {&s->rhovx, &s->rhovy, &s->rhovz};
            A3 *mp[3] = {&s->rhovxp, &s->rhovyp, &s->rhovzp};
            for (int c = 0; c < 3; ++c) {
                if (st == 2 && nst == 2) m[c]->d[q] = (mp[c]->d[q] + m[c]->d[q]) / 2.;
                else if (st == 2) m[c]->d[q] = (3. * mp[c]->d[q] + m[c]->d[q]) / 4.;
                else if (st == 3) m[c]->d[q] = (mp[c]->d[q] + 2. * m[c]->d[q]) / 3.;
            }
            s->vx.d[q] = s->rhovx.d[q] / s->rho.d[q];
            s->vy.d[q] = s->rhovy.d[q] / s->rho.d[q];
            s->vz.d[q] = s->rhovz.d[q] / s->rho.d[q];
        });
        advance_bijk(s);
        if (st >= 2) {
            A3 *b[3] = {&s->bi, &s->bj, &s->bk};
            A3 *bp[3] = {&s->bip, &s->bjp, &s->bkp};
            for (int c = 0; c < 3; ++c) {
                size_t n = a3_size(b[c]);
                for (size_t q = 0; q < n; ++q) {
                    if (nst == 2) b[c]->d[q] = (bp[c]->d[q] + b[c]->d[q]) / 2.;
                    else if (st == 2) b[c]->d[q] = (3. * bp[c]->d[q] + b[c]->d[q]) / 4.;
                    else b[c]->d[q] = (bp[c]->d[q] + 2. * b[c]->d[q]) / 3.;
                }
            }
        }
    }
}

/* TVD RK2 / RK3 (mhd3d3.f90:3648-3792); nst = number of stages */
static double step_rk(orc_state *s, int nst) {
    double tcfl = orc_get_dt(s);
    s->tau = s->tau0;
    advance_previous(s);
    for (int st = 1; st <= nst; ++st) {
        get_conserve(s);
        eijk_lfm(s, &s->rho, &s->vx, &s->vy, &s->vz, &s->bx, &s->by, &s->bz, &s->bi, &s->bj, &s->bk);
        flux_stress(s, &s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz);
        rk_update(s, nst, st);
        orc_fill(s);
        orc_bxyz(s);
    }
    return tcfl;
}

/* advance_predictor_ab (mhd3d3.f90:3821-3840) and the tau0=tau of step_ab (3616) */
static void ab_predict(orc_state *s) {
    const double c = s->tau / s->tau0 / 2.;
    A3 *x[11] = {&s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz, &s->bi, &s->bj, &s->bk};
    A3 *xp[11] = {&s->rhop, &s->vxp, &s->vyp, &s->vzp, &s->pp, &s->bxp, &s->byp, &s->bzp, &s->bip, &s->bjp, &s->bkp};
    A3 *xh[11] = {&s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh, &s->bih, &s->bjh, &s->bkh};
    for (int a = 0; a < 11; ++a) {
        size_t n = a3_size(x[a]);
        for (size_t q = 0; q < n; ++q) xh[a]->d[q] = x[a]->d[q] + c * (x[a]->d[q] - xp[a]->d[q]);
    }
    advance_previous(s);
    s->tau0 = s->tau;
}

/* the Adams-Bashforth update after the flux (mhd3d3.f90:3622-3627):
 * corrector1 with the pressure, corrector2 with the velocity, advance_bijk */
static void ab_update(orc_state *s) {
    const double gam = s->p.gam;
    corrector1(s);
    FOR_INTERIOR(s, { s->pr.d[q] = prs(gam, s->eng.d[q], s->rhovx.d[q], s->rhovy.d[q], s->rhovz.d[q], s->rho.d[q]); });
    corrector2(s);
    FOR_INTERIOR(s, {
        s->vx.d[q] = s->rhovx.d[q] / s->rho.d[q];
        s->vy.d[q] = s->rhovy.d[q] / s->rho.d[q];
        s->vz.d[q] = s->rhovz.d[q] / s->rho.d[q];
    });
    advance_bijk(s);
}

/* Adams-Bashforth predictor-corrector (mhd3d3.f90:3602-3644) */
static void step_ab(orc_state *s) {
    orc_get_dt(s);
    get_conserve(s);
    ab_predict(s);
    eijk_lfm(s, &s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->bxh, &s->byh, &s->bzh, &s->bih, &s->bjh, &s->bkh);
    flux_stress(s, &s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh);
    ab_update(s);
    orc_fill(s);
    orc_bxyz(s);
}

int orc_step(orc_state *s, int nsteps, double *tau, double *tau_cfl) {
    double tcfl = 0.;
    if (s->p.stepop != 1 && tau) s->tau0 = *tau; /* RK: fixed deck dt (read_input: tau0=tau) */
    for (int n = 0; n < nsteps; ++n) {
        if (s->p.stepop == 1) {
            step_ab(s);
            tcfl = s->tau;
        } else if (s->p.stepop == 2 || s->p.stepop == 3) {
            tcfl = step_rk(s, s->p.stepop == 2 ? 2 : 3);
        } else {
            return 1; /* 'Undefined time adavnce option!' (mhd3d3.f90:3596) */
        }
    }
    if (tau) *tau = s->tau;
    if (tau_cfl) *tau_cfl = tcfl;
    return 0;
}

/* --------------------------------------------------------------------- */
/* lifecycle                                                              */
/* --------------------------------------------------------------------- */
orc_state *orc_create(const orc_params *p) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    int nx = p->nx0 + 2 * p->ng, ny = p->ny0 + 2 * p->ng, nz = p->nz0 + 2 * p->ng;
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->dx = (double *)calloc(nx, sizeof(double));
    s->dy = (double *)calloc(ny, sizeof(double));
    s->dz = (double *)calloc(nz, sizeof(double));
    A3 *cells[] = {&s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz, &s->rhovx, &s->rhovy,
                   &s->rhovz, &s->eng, &s->rho0, &s->rhovx0, &s->rhovy0, &s->rhovz0, &s->eng0, &s->rhop,
                   &s->vxp, &s->vyp, &s->vzp, &s->pp, &s->bxp, &s->byp, &s->bzp, &s->rhoh, &s->vxh,
                   &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh, &s->rhovxp, &s->rhovyp,
                   &s->rhovzp, &s->rhovxh, &s->rhovyh, &s->rhovzh, &s->engp};
    for (size_t q = 0; q < sizeof(cells) / sizeof(cells[0]); ++q) *cells[q] = a3_new(nx, ny, nz);
    s->bi = a3_new(nx + 1, ny, nz);
    s->bj = a3_new(nx, ny + 1, nz);
    s->bk = a3_new(nx, ny, nz + 1);
    s->bip = a3_new(nx + 1, ny, nz);
    s->bjp = a3_new(nx, ny + 1, nz);
    s->bkp = a3_new(nx, ny, nz + 1);
    s->bih = a3_new(nx + 1, ny, nz);
    s->bjh = a3_new(nx, ny + 1, nz);
    s->bkh = a3_new(nx, ny, nz + 1);
    s->ei = a3_new(p->nx0, p->ny0 + 1, p->nz0 + 1);
    s->ej = a3_new(p->nx0 + 1, p->ny0, p->nz0 + 1);
    s->ek = a3_new(p->nx0 + 1, p->ny0 + 1, p->nz0);
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 8; ++q)
            s->flux[8 * d + q] = a3_new(p->nx0 + (d == 0), p->ny0 + (d == 1), p->nz0 + (d == 2));
    s->tau = s->tau0 = 0.;
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    A3 *all[] = {&s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz, &s->rhovx, &s->rhovy,
                 &s->rhovz, &s->eng, &s->rho0, &s->rhovx0, &s->rhovy0, &s->rhovz0, &s->eng0, &s->rhop,
                 &s->vxp, &s->vyp, &s->vzp, &s->pp, &s->bxp, &s->byp, &s->bzp, &s->rhoh, &s->vxh,
                 &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh, &s->rhovxp, &s->rhovyp,
                 &s->rhovzp, &s->rhovxh, &s->rhovyh, &s->rhovzh, &s->engp, &s->bi, &s->bj, &s->bk,
                 &s->bip, &s->bjp, &s->bkp, &s->bih, &s->bjh, &s->bkh, &s->ei, &s->ej, &s->ek};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) a3_free(all[q]);
    for (int q = 0; q < ORC_NFLUX; ++q) a3_free(&s->flux[q]);
    free(s->dx);
    free(s->dy);
    free(s->dz);
    free(s->sur_w);
    free(s);
}

void orc_set_geometry(orc_state *s, const double *dx, const double *dy, const double *dz) {
    memcpy(s->dx, dx, s->nx * sizeof(double));
    memcpy(s->dy, dy, s->ny * sizeof(double));
    memcpy(s->dz, dz, s->nz * sizeof(double));
}

void orc_set_state(orc_state *s, const double *rho, const double *vx, const double *vy, const double *vz,
                   const double *p, const double *bi, const double *bj, const double *bk) {
    size_t n = a3_size(&s->rho);
    memcpy(s->rho.d, rho, n * sizeof(double));
    memcpy(s->vx.d, vx, n * sizeof(double));
    memcpy(s->vy.d, vy, n * sizeof(double));
    memcpy(s->vz.d, vz, n * sizeof(double));
    memcpy(s->pr.d, p, n * sizeof(double));
    memcpy(s->bi.d, bi, a3_size(&s->bi) * sizeof(double));
    memcpy(s->bj.d, bj, a3_size(&s->bj) * sizeof(double));
    memcpy(s->bk.d, bk, a3_size(&s->bk) * sizeof(double));
}

void orc_get_state(const orc_state *s, double *rho, double *vx, double *vy, double *vz, double *p, double *bi,
                   double *bj, double *bk, double *bx, double *by, double *bz) {
    size_t n = a3_size(&s->rho);
    if (rho) memcpy(rho, s->rho.d, n * sizeof(double));
    if (vx) memcpy(vx, s->vx.d, n * sizeof(double));
    if (vy) memcpy(vy, s->vy.d, n * sizeof(double));
    if (vz) memcpy(vz, s->vz.d, n * sizeof(double));
    if (p) memcpy(p, s->pr.d, n * sizeof(double));
    if (bi) memcpy(bi, s->bi.d, a3_size(&s->bi) * sizeof(double));
    if (bj) memcpy(bj, s->bj.d, a3_size(&s->bj) * sizeof(double));
    if (bk) memcpy(bk, s->bk.d, a3_size(&s->bk) * sizeof(double));
    if (bx) memcpy(bx, s->bx.d, n * sizeof(double));
    if (by) memcpy(by, s->by.d, n * sizeof(double));
    if (bz) memcpy(bz, s->bz.d, n * sizeof(double));
}

/* initialization (mhd3d3.f90:3327-3377) after init_ic */
void orc_initialize(orc_state *s, double *tau) {
    s->tau = s->tau0 = tau ? *tau : 0.; /* read_input: tau0=tau */
    orc_fill(s);
    orc_bxyz(s);
    if (s->p.stepop == 1) {
        orc_get_dt(s);
        s->tau = 3.e-4; /* hard-coded AB start (mhd3d3.f90:3359) */
        s->tau0 = s->tau;
        a3_copy(&s->rhop, &s->rho);
        a3_copy(&s->vxp, &s->vx);
        a3_copy(&s->vyp, &s->vy);
        a3_copy(&s->vzp, &s->vz);
        a3_copy(&s->pp, &s->pr);
        a3_copy(&s->bxp, &s->bx);
        a3_copy(&s->byp, &s->by);
        a3_copy(&s->bzp, &s->bz);
        a3_copy(&s->bip, &s->bi);
        a3_copy(&s->bjp, &s->bj);
        a3_copy(&s->bkp, &s->bk);
    }
    if (tau) *tau = s->tau;
}

void orc_flux_stress(orc_state *s) {
    flux_stress(s, &s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz);
}

void orc_eijk(orc_state *s) {
    eijk_lfm(s, &s->rho, &s->vx, &s->vy, &s->vz, &s->bx, &s->by, &s->bz, &s->bi, &s->bj, &s->bk);
}

void orc_set_tau(orc_state *s, double tau, double tau0) {
    s->tau = tau;
    s->tau0 = tau0;
}

void orc_get_tau(const orc_state *s, double *tau, double *tau0) {
    if (tau) *tau = s->tau;
    if (tau0) *tau0 = s->tau0;
}

int orc_predict_ab(orc_state *s) {
    if (s->p.stepop != 1) return 1;
    ab_predict(s);
    return 0;
}

void orc_flux_stress_state(orc_state *s, int which) {
    if (which == 1)
        flux_stress(s, &s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh);
    else
        orc_flux_stress(s);
}

void orc_eijk_state(orc_state *s, int which) {
    if (which == 1)
        eijk_lfm(s, &s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->bxh, &s->byh, &s->bzh, &s->bih, &s->bjh, &s->bkh);
    else
        orc_eijk(s);
}

/* --------------------------------------------------------------------- */
/* get_eijk_Yee (mhd3d3.f90:1654-1752) with get_v_center (2562-2592)       */
/* --------------------------------------------------------------------- */

/* reconstruct_3d of a whole array into (l1, l2, l3) (mhd3d3.f90:833-860) */
static void recon_arr(const orc_params *p, const A3 *f, int dir, A3 *fl, A3 *fr) {
    orc_reconstruct(p, f->d, f->n1, f->n2, f->n3, fl->n1, fl->n2, fl->n3, dir, fl->d, fr->d);
}

static void eijk_yee(orc_state *s, const A3 *lrho, const A3 *lvx, const A3 *lvy, const A3 *lvz, const A3 *lp,
                     const A3 *lbx, const A3 *lby, const A3 *lbz) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, ng = s->p.ng;
    const int nx0 = s->p.nx0, ny0 = s->p.ny0, nz0 = s->p.nz0;
    const size_t n = (size_t)nx * ny * nz;
    /* split fields e[a][pp, np, pn, nn] (1669-1691): first split the
     * electric field: bsq, lam, get_v_center's moments per axis, then
     * e.g. ezpp = lbx*vx0p*vy1p - lby*vx1p*vy0p */
    A3 e[3][4];
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 4; ++q) e[a][q] = a3_new(nx, ny, nz);
    const A3 *vv[3] = {lvx, lvy, lvz}, *bb[3] = {lbx, lby, lbz};
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < n; ++t) {
        double bsq = lbx->d[t] * lbx->d[t] + lby->d[t] * lby->d[t] + lbz->d[t] * lbz->d[t];
        double lam = lrho->d[t] / (2. * lp->d[t] + bsq);
        double m0[3][2], m1[3][2];
        for (int a = 0; a < 3; ++a) vface_pt(lam, vv[a]->d[t], &m0[a][0], &m0[a][1], &m1[a][0], &m1[a][1]);
        for (int a = 0; a < 3; ++a) {
            int b = (a + 1) % 3, c = (a + 2) % 3;
            for (int q = 0; q < 4; ++q) {
                int s1 = q & 1, s2 = q >> 1; /* pp, np, pn, nn */
                e[a][q].d[t] = bb[b]->d[t] * m0[b][s1] * m1[c][s2] - bb[c]->d[t] * m1[b][s1] * m0[c][s2];
            }
        }
    }
    /* per edge direction: (first transverse direction, its extents),
     * (second, its extents) -- ei: y then z (1696-1714), ej: z then x
     * (1722-1739), ek: x then y (1747-1764); 5 = y, 6 = z, 4 = x untruncated */
    const int dir1[3] = {5, 6, 4}, dir2[3] = {6, 4, 5};
    const int l1[3][3] = {{nx, ny0 + 1, nz}, {nx, ny, nz0 + 1}, {nx0 + 1, ny, nz}};
    const int l2[3][3] = {{nx, ny0 + 1, nz0 + 1}, {nx0 + 1, ny, nz0 + 1}, {nx0 + 1, ny0 + 1, nz}};
    A3 *out[3] = {&s->ei, &s->ej, &s->ek};
    for (int a = 0; a < 3; ++a) {
        A3 q[4], d1 = a3_new(l1[a][0], l1[a][1], l1[a][2]);
        for (int k = 0; k < 4; ++k) q[k] = a3_new(l1[a][0], l1[a][1], l1[a][2]);
        /* qepp, qepn: left states; qenn, qenp: right states */
        recon_arr(&s->p, &e[a][0], dir1[a], &q[0], &d1);
        recon_arr(&s->p, &e[a][2], dir1[a], &q[2], &d1);
        recon_arr(&s->p, &e[a][3], dir1[a], &d1, &q[3]);
        recon_arr(&s->p, &e[a][1], dir1[a], &d1, &q[1]);
        A3 epp = a3_new(l2[a][0], l2[a][1], l2[a][2]), enp = a3_new(l2[a][0], l2[a][1], l2[a][2]);
        A3 enn = a3_new(l2[a][0], l2[a][1], l2[a][2]), epn = a3_new(l2[a][0], l2[a][1], l2[a][2]);
        A3 d2 = a3_new(l2[a][0], l2[a][1], l2[a][2]);
        recon_arr(&s->p, &q[0], dir2[a], &epp, &d2);  /* epp: left of qepp */
        recon_arr(&s->p, &q[1], dir2[a], &enp, &d2);  /* enp: left of qenp */
        recon_arr(&s->p, &q[3], dir2[a], &d2, &enn);  /* enn: right of qenn */
        recon_arr(&s->p, &q[2], dir2[a], &d2, &epn);  /* epn: right of qepn */
        /* ei(i,:,:)=epp(ng+i,:,:)+enp(ng+i,:,:)+epn(ng+i,:,:)+enn(ng+i,:,:) (1717) and kin */
        A3 *o = out[a];
        for (int k = 1; k <= o->n3; ++k)
            for (int j = 1; j <= o->n2; ++j)
                for (int i = 1; i <= o->n1; ++i) {
                    int I = i + (a == 0 ? ng : 0), J = j + (a == 1 ? ng : 0), K = k + (a == 2 ? ng : 0);
                    AT(*o, i, j, k) = AT(epp, I, J, K) + AT(enp, I, J, K) + AT(epn, I, J, K) + AT(enn, I, J, K);
                }
        for (int k = 0; k < 4; ++k) a3_free(&q[k]);
        a3_free(&d1);
        a3_free(&epp);
        a3_free(&enp);
        a3_free(&enn);
        a3_free(&epn);
        a3_free(&d2);
    }
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 4; ++q) a3_free(&e[a][q]);
}

void orc_eijk_yee_state(orc_state *s, int which) {
    if (which == 1)
        eijk_yee(s, &s->rhoh, &s->vxh, &s->vyh, &s->vzh, &s->ph, &s->bxh, &s->byh, &s->bzh);
    else
        eijk_yee(s, &s->rho, &s->vx, &s->vy, &s->vz, &s->pr, &s->bx, &s->by, &s->bz);
}

/* get_conserve and the stage update from the held fluxes and EMFs: stage 0
 * the Adams-Bashforth corrector, 1..nst a TVD RK stage (stage 1 after the
 * advance_previous that step_rk2 takes just before it, mhd3d3.f90:3664) */
int orc_advance(orc_state *s, int stage) {
    const int nst = s->p.stepop == 1 ? 0 : s->p.stepop;
    if (nst == 0 ? stage != 0 : (stage < 1 || stage > nst)) return 1;
    if (stage == 1) advance_previous(s);
    get_conserve(s);
    if (nst == 0) ab_update(s);
    else rk_update(s, nst, stage);
    return 0;
}

void orc_set_ops(orc_state *s, const double *ei, const double *ej, const double *ek, const double *flux_all) {
    if (ei) memcpy(s->ei.d, ei, a3_size(&s->ei) * sizeof(double));
    if (ej) memcpy(s->ej.d, ej, a3_size(&s->ej) * sizeof(double));
    if (ek) memcpy(s->ek.d, ek, a3_size(&s->ek) * sizeof(double));
    if (flux_all) {
        size_t off = 0;
        for (int q = 0; q < ORC_NFLUX; ++q) {
            size_t n = a3_size(&s->flux[q]);
            memcpy(s->flux[q].d, flux_all + off, n * sizeof(double));
            off += n;
        }
    }
}

void orc_get_ops(const orc_state *s, double *ei, double *ej, double *ek, double *flux_all) {
    if (ei) memcpy(ei, s->ei.d, a3_size(&s->ei) * sizeof(double));
    if (ej) memcpy(ej, s->ej.d, a3_size(&s->ej) * sizeof(double));
    if (ek) memcpy(ek, s->ek.d, a3_size(&s->ek) * sizeof(double));
    if (flux_all) {
        size_t off = 0;
        for (int q = 0; q < ORC_NFLUX; ++q) {
            size_t n = a3_size(&s->flux[q]);
            memcpy(flux_all + off, s->flux[q].d, n * sizeof(double));
            off += n;
        }
    }
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* --------------------------------------------------------------------- */
/* learned reconstruction: CPU restatement of the net (checker for the   */
/* GPU training and inference kernels).  No reference counterpart.       */
/* y = lin(x) + gate(x) * MLP(x), per output of variable q:             */
/* lin = (-x0 + 7x1 + 7x2 - x3)/12, gate = max |x(k-1) - 2x(k) + x(k+1)| */
/* of q's window                                                          */
/* layout: W1[32][40] b1[32] W2[32][32] b2[32] W3[16][32] b3[16]         */
/* activation g(z) = z*sigmoid(1.702 z); loss = mean squared error over  */
/* the 16 outputs of the n samples (56 floats: 40 inputs, 16 targets)    */
/* --------------------------------------------------------------------- */
#define MNI 40
#define MNH 32
#define MNO 16
#define MOFF_B1 (MNH * MNI)
#define MOFF_W2 (MOFF_B1 + MNH)
#define MOFF_B2 (MOFF_W2 + MNH * MNH)
#define MOFF_W3 (MOFF_B2 + MNH)
#define MOFF_B3 (MOFF_W3 + MNO * MNH)
#define MNPARAM (MOFF_B3 + MNO)

static double mact(double z) { return z / (1. + exp(-1.702 * z)); }
/* fixed linear part: 4th-order interpolation of the window of output o */
static double mlin(const float *x, int o) {
    const float *w = x + 5 * (o & 7);
    return (-(double)w[0] + 7. * w[1] + 7. * w[2] - w[3]) / 12.;
}
/* gate of output o: largest |second difference| of its variable's window,
 * so the learned correction vanishes on linear data */
static double mgate(const float *x, int o) {
    const float *w = x + 5 * (o & 7);
    double g = 0.;
    for (int a = 1; a < 4; ++a) g = fmax(g, fabs((double)w[a - 1] - 2. * w[a] + w[a + 1]));
    return g;
}
static double mact_grad(double z) {
    double s = 1. / (1. + exp(-1.702 * z));
    return s + 1.702 * z * s * (1. - s);
}

/* forward of n rows of 40 inputs; h_round != 0 rounds the hidden
 * activations to bf16 (round to nearest even) like the MFMA inference */
static double bf16r(double v) {
    float f = (float)v;
    unsigned int u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

void orc_mlp_forward(const float *w, const float *x, long n, int h_round, float *y) {
    for (long s = 0; s < n; ++s) {
        const float *xi = x + s * MNI;
        double a1[MNH], a2[MNH];
        for (int j = 0; j < MNH; ++j) {
            double acc = 0.;
            for (int a = 0; a < MNI; ++a) acc += (double)w[j * MNI + a] * xi[a];
            acc += w[MOFF_B1 + j];
            a1[j] = h_round ? bf16r(mact(acc)) : mact(acc);
        }
        for (int j = 0; j < MNH; ++j) {
            double acc = 0.;
            for (int a = 0; a < MNH; ++a) acc += (double)w[MOFF_W2 + j * MNH + a] * a1[a];
            acc += w[MOFF_B2 + j];
            a2[j] = h_round ? bf16r(mact(acc)) : mact(acc);
        }
        for (int o = 0; o < MNO; ++o) {
            double acc = 0.;
            for (int a = 0; a < MNH; ++a) acc += (double)w[MOFF_W3 + o * MNH + a] * a2[a];
            y[s * MNO + o] = (float)(mgate(xi, o) * (acc + w[MOFF_B3 + o]) + mlin(xi, o));
        }
    }
}

#define SUR_ROUGH_MIN 1e-3

/* the learned face states of one face (the solver's use of the net):
 * windows of cells j-2..j+2 centred on c = (f(j-1)+f(j))/2, divided by
 * max(range, 1e-3 * physical scale), net y = lin + MLP, clamped to the two
 * cells' range; returns 0 (PDM fallback) when the largest normalised second
 * difference exceeds theta or a density/pressure state is not positive.
 * With sur_round the weights, inputs and hidden activations are rounded to
 * bf16 as in the MFMA inference. */
static int sur_face(const orc_state *s, const A3 *const *f, int dir, int i, int j, int k, double *lv,
                    double *rv) {
    const int ng = s->p.ng;
    double w5[8][5], c[8], m[8], sc[8];
    float xf[MNI];
    for (int v = 0; v < 8; ++v) {
        for (int a = 0; a < 5; ++a) w5[v][a] = gat(f[v], ng, dir, i, j, k, a - 2, 0, 0, 0);
        c[v] = .5 * (w5[v][1] + w5[v][2]);
        m[v] = 0.;
        for (int a = 0; a < 5; ++a) m[v] = fmax(m[v], fabs(w5[v][a] - c[v]));
    }
    {
        const double r = fabs(c[0]), pp = fabs(c[4]);
        const double vs = fmax(fmax(fabs(c[1]), fabs(c[2])), fabs(c[3])) + sqrt(pp / fmax(r, 1e-30));
        const double bs = fmax(fmax(fabs(c[5]), fabs(c[6])), fabs(c[7])) + sqrt(pp);
        sc[0] = r, sc[1] = sc[2] = sc[3] = vs, sc[4] = pp, sc[5] = sc[6] = sc[7] = bs;
    }
    double ss[8], rough = 0.;
    for (int v = 0; v < 8; ++v) {
        ss[v] = fmax(fmax(m[v], 1e-3 * sc[v]), 1e-30);
        for (int a = 0; a < 5; ++a) xf[5 * v + a] = (float)((w5[v][a] - c[v]) / ss[v]);
        for (int a = 1; a < 4; ++a)
            rough = fmax(rough, fabs((double)xf[5 * v + a - 1] - 2. * xf[5 * v + a] + xf[5 * v + a + 1]));
    }
    float y[MNO];
    if (s->sur_round) {
        float xr[MNI];
        for (int a = 0; a < MNI; ++a) xr[a] = (float)bf16r(xf[a]);
        orc_mlp_forward(s->sur_w, xr, 1, 1, y);
        /* the fixed linear part and the gate use the unrounded inputs */
        for (int o = 0; o < MNO; ++o) {
            const double gr = mgate(xr, o), gf = mgate(xf, o);
            const double corr = gr > 0. ? (y[o] - mlin(xr, o)) / gr : 0.;
            y[o] = (float)(mlin(xf, o) + gf * corr);
        }
    } else {
        orc_mlp_forward(s->sur_w, xf, 1, 0, y);
    }
    /* windows curved by at most SUR_ROUGH_MIN take the linear part alone */
    if (!(rough > SUR_ROUGH_MIN))
        for (int o = 0; o < MNO; ++o) y[o] = (float)mlin(xf, o);
    for (int v = 0; v < 8; ++v) {
        const double lo = fmin(xf[5 * v + 1], xf[5 * v + 2]), hi = fmax(xf[5 * v + 1], xf[5 * v + 2]);
        lv[v] = c[v] + ss[v] * fmin(hi, fmax(lo, (double)y[v]));
        rv[v] = c[v] + ss[v] * fmin(hi, fmax(lo, (double)y[8 + v]));
    }
    return rough <= s->sur_theta && lv[0] > 0. && rv[0] > 0. && lv[4] > 0. && rv[4] > 0.;
}

void orc_set_surrogate(orc_state *s, const float *w, double theta, int bf16_round) {
    free(s->sur_w);
    s->sur_w = NULL;
    if (!w) return;
    s->sur_w = (float *)malloc(MNPARAM * sizeof(float));
    memcpy(s->sur_w, w, MNPARAM * sizeof(float));
    if (bf16_round)
        for (int q = 0; q < MNPARAM; ++q) s->sur_w[q] = (float)bf16r(s->sur_w[q]);
    s->sur_theta = theta;
    s->sur_round = bf16_round;
}

/* loss and gradient with respect to the parameters (double accumulation) */
double orc_mlp_loss_grad(const float *w, const float *samples, long n, double *grad) {
    double loss = 0.;
    for (int q = 0; q < MNPARAM; ++q) grad[q] = 0.;
    const double inv = 1. / ((double)n * MNO);
    for (long s = 0; s < n; ++s) {
        const float *x = samples + s * (MNI + MNO), *t = x + MNI;
        double z1[MNH], a1[MNH], z2[MNH], a2[MNH], dy[MNO], dz2[MNH], dz1[MNH];
        for (int j = 0; j < MNH; ++j) {
            double acc = w[MOFF_B1 + j];
            for (int a = 0; a < MNI; ++a) acc += (double)w[j * MNI + a] * x[a];
            z1[j] = acc;
            a1[j] = mact(acc);
        }
        for (int j = 0; j < MNH; ++j) {
            double acc = w[MOFF_B2 + j];
            for (int a = 0; a < MNH; ++a) acc += (double)w[MOFF_W2 + j * MNH + a] * a1[a];
            z2[j] = acc;
            a2[j] = mact(acc);
        }
        for (int o = 0; o < MNO; ++o) {
            double acc = w[MOFF_B3 + o];
            for (int a = 0; a < MNH; ++a) acc += (double)w[MOFF_W3 + o * MNH + a] * a2[a];
            const double gt = mgate(x, o);
            double e = gt * acc + mlin(x, o) - t[o];
            loss += e * e;
            dy[o] = 2. * e * gt * inv;
        }
        for (int a = 0; a < MNH; ++a) {
            double acc = 0.;
            for (int o = 0; o < MNO; ++o) acc += (double)w[MOFF_W3 + o * MNH + a] * dy[o];
            dz2[a] = acc * mact_grad(z2[a]);
        }
        for (int a = 0; a < MNH; ++a) {
            double acc = 0.;
            for (int j = 0; j < MNH; ++j) acc += (double)w[MOFF_W2 + j * MNH + a] * dz2[j];
            dz1[a] = acc * mact_grad(z1[a]);
        }
        for (int j = 0; j < MNH; ++j) {
            for (int a = 0; a < MNI; ++a) grad[j * MNI + a] += dz1[j] * x[a];
            grad[MOFF_B1 + j] += dz1[j];
            for (int a = 0; a < MNH; ++a) grad[MOFF_W2 + j * MNH + a] += dz2[j] * a1[a];
            grad[MOFF_B2 + j] += dz2[j];
        }
        for (int o = 0; o < MNO; ++o) {
            for (int a = 0; a < MNH; ++a) grad[MOFF_W3 + o * MNH + a] += dy[o] * a2[a];
            grad[MOFF_B3 + o] += dy[o];
        }
    }
    return loss * inv;
}
