/*
 * cns_oracle.c -- CPU restatement of the CNS_NoSpec time step (TEST INFRASTRUCTURE).
 *
 * Only tests/, the smoke check and bench.py's cpu_baseline leg use this file.
 * The HIP library never links it and has no CPU fallback.
 *
 * Built with `gcc -O2 -ffp-contract=off` (oracle/Makefile).  Each routine keeps
 * the reference's operation order (parenthesisation, left-to-right sums, the
 * x-then-y-then-z accumulation of hypterm, `/dx` in courno but `*dxinv` in the
 * stencils), so that its output is bit-identical to the reference cpu/ path
 * compiled with g++ -O2 on x86-64 (no FMA there).  The only deliberate change:
 * the Courant maxima start at 0 instead of an uninitialised read
 * (cpu/ctoprim.cpp:21,57-59); every Courant number is positive, so the result
 * is the same closed-form maximum.
 *
 * The per-cell loops run on OpenMP threads (orc_set_threads); every cell is
 * computed by the same expression from the same inputs whatever the thread, so
 * the results do not depend on the thread count.  The Courant maximum stays a
 * serial loop (its order decides which of several NaNs survives).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "cns_oracle.h"

/* cpu/header.h:14-45 */
#define NG 4
#define NC 5
enum { irho = 0, imx, imy, imz, iene };
enum { qrho = 0, qu, qv, qw, qpres, qtemp };

/* cpu/header.h:88-91 (first-derivative weights) */
static const double ALP = 0.8E0;
static const double BET = -0.2E0;
static const double GAM = 4.0E0 / 105.0E0;
static const double DEL = -1.0E0 / 280.0E0;

#define PAD16(x) (((((x) - 1) >> 4) + 1) << 4) /* gpu/header.h:14 */
#define MAX(x, y) ((x > y) ? (x) : (y))         /* cpu/header.h:11 */

typedef struct {
    int n[3], ng, dimg[3];
    long xg, pg, cg; /* ghosted pitches */
    long x, p, c;    /* interior pitches */
} layout_t;

static int make_layout(const int lo[3], const int hi[3], int ng, layout_t *L)
{
    int d;
    if (ng != NG) return -1;
    L->ng = ng;
    for (d = 0; d < 3; d++) {
        L->n[d] = hi[d] - lo[d] + 1;
        if (L->n[d] < 1 || lo[d] != ng) return -1;
        L->dimg[d] = L->n[d] + 2 * ng;
    }
    L->xg = PAD16(L->dimg[0]);
    L->pg = L->xg * L->dimg[1];
    L->cg = L->pg * L->dimg[2];
    L->x = PAD16(L->n[0]);
    L->p = L->x * L->n[1];
    L->c = L->p * L->n[2];
    return 0;
}

#define GIDX(L, k, j, i) ((long)(k) * (L)->pg + (long)(j) * (L)->xg + (long)(i))
#define IIDX(L, k, j, i) ((long)(k) * (L)->p + (long)(j) * (L)->x + (long)(i))

/* ------------------------------------------------------------------ fill */
/* cpu/helper_functions.cpp:226-262; SoA order as gpu/util_cpu.cpp:127-169. */
int orc_fill_boundary(double *u, const int dim[3], const int dim_g[3])
{
    int lo[3], hi[3], d, i, j, k, l;
    layout_t L;
    for (d = 0; d < 3; d++) {
        lo[d] = NG;
        hi[d] = dim[d] - 1 + NG;
        if (dim_g[d] != dim[d] + 2 * NG) return -1;
    }
    if (make_layout(lo, hi, NG, &L)) return -1;
    for (l = 0; l < NC; l++) {
        double *U = u + (long)l * L.cg;
#pragma omp parallel for collapse(2) private(i)
        for (k = 0; k < NG; k++)
            for (j = NG; j < dim[1] + NG; j++)
                for (i = NG; i < dim[0] + NG; i++) {
                    U[GIDX(&L, k, j, i)] = U[GIDX(&L, k + dim[2], j, i)];
                    U[GIDX(&L, k + dim[2] + NG, j, i)] = U[GIDX(&L, k + NG, j, i)];
                }
    }
    for (l = 0; l < NC; l++) {
        double *U = u + (long)l * L.cg;
#pragma omp parallel for collapse(2) private(i)
        for (k = 0; k < dim_g[2]; k++)
            for (j = 0; j < NG; j++)
                for (i = NG; i < dim[0] + NG; i++) {
                    U[GIDX(&L, k, j, i)] = U[GIDX(&L, k, j + dim[1], i)];
                    U[GIDX(&L, k, j + dim[1] + NG, i)] = U[GIDX(&L, k, j + NG, i)];
                }
    }
    for (l = 0; l < NC; l++) {
        double *U = u + (long)l * L.cg;
#pragma omp parallel for collapse(2) private(i)
        for (k = 0; k < dim_g[2]; k++)
            for (j = 0; j < dim_g[1]; j++)
                for (i = 0; i < NG; i++) {
                    U[GIDX(&L, k, j, i)] = U[GIDX(&L, k, j, i + dim[0])];
                    U[GIDX(&L, k, j, i + dim[0] + NG)] = U[GIDX(&L, k, j, i + NG)];
                }
    }
    return 0;
}

/* --------------------------------------------------------------- ctoprim */
/* cpu/ctoprim.cpp:11-70 (with courno) and :72-105 (without). */
int orc_ctoprim(const int lo[3], const int hi[3], const double *u, double *q,
                const double dx[3], int ng, double *courno)
{
    const double GAMMA = 1.4E0;
    const double CV = 8.3333333333E6;
    layout_t L;
    int i, j, k;
    long cg;
    if (make_layout(lo, hi, ng, &L)) return -1;
    cg = L.cg;

#pragma omp parallel for collapse(2) private(i)
    for (k = 0; k < L.dimg[2]; k++)
        for (j = 0; j < L.dimg[1]; j++)
            for (i = 0; i < L.dimg[0]; i++) {
                long g = GIDX(&L, k, j, i);
                double rhoinv, eint;
                rhoinv = 1.0E0 / u[g];
                q[g] = u[g];
                q[g + cg] = u[g + cg] * rhoinv;
                q[g + 2 * cg] = u[g + 2 * cg] * rhoinv;
                q[g + 3 * cg] = u[g + 3 * cg] * rhoinv;
                eint = u[g + 4 * cg] * rhoinv
                       - 0.5E0 * ((q[g + cg] * q[g + cg]) + (q[g + 2 * cg] * q[g + 2 * cg])
                                  + (q[g + 3 * cg] * q[g + 3 * cg]));
                q[g + 4 * cg] = (GAMMA - 1.0E0) * eint * u[g];
                q[g + 5 * cg] = eint / CV;
            }

    if (courno) {
        double courmx = 0.0, courmy = 0.0, courmz = 0.0;
        for (k = lo[2]; k <= hi[2]; k++)
            for (j = lo[1]; j <= hi[1]; j++)
                for (i = lo[0]; i <= hi[0]; i++) {
                    long g = GIDX(&L, k, j, i);
                    double c, courx, coury, courz;
                    c = sqrt(GAMMA * q[g + 4 * cg] / q[g]);
                    courx = (c + fabs(q[g + cg])) / dx[0];
                    coury = (c + fabs(q[g + 2 * cg])) / dx[1];
                    courz = (c + fabs(q[g + 3 * cg])) / dx[2];
                    courmx = MAX(courmx, courx);
                    courmy = MAX(courmy, coury);
                    courmz = MAX(courmz, courz);
                }
        *courno = MAX(MAX(courmx, courmy), MAX(courmz, *courno));
    }
    return 0;
}

/* --------------------------------------------------------------- hypterm */
/* Sums A[l] of one direction of cpu/hypterm.cpp (x: :33-90, y: :101-147, z: :158-204),
 * before the multiplication by dxinv.  s = stride of the direction, qn = normal
 * velocity, mn = normal momentum. */
static void hyp_sums(const double *cons, const double *q, long g, long s, long cg,
                     int qn, int mn, double A[5])
{
#define C(l, m) cons[(long)(l) * cg + g + (m) * s]
#define QN(m) q[(long)qn * cg + g + (m) * s]
#define P(m) q[(long)qpres * cg + g + (m) * s]
    double unp1 = QN(1), unp2 = QN(2), unp3 = QN(3), unp4 = QN(4);
    double unm1 = QN(-1), unm2 = QN(-2), unm3 = QN(-3), unm4 = QN(-4);
    int l;

    A[irho] = ALP * (C(mn, 1) - C(mn, -1))
              + BET * (C(mn, 2) - C(mn, -2))
              + GAM * (C(mn, 3) - C(mn, -3))
              + DEL * (C(mn, 4) - C(mn, -4));
    for (l = imx; l <= imz; l++) {
        if (l == mn)
            A[l] = ALP * (C(l, 1) * unp1 - C(l, -1) * unm1 + (P(1) - P(-1)))
                   + BET * (C(l, 2) * unp2 - C(l, -2) * unm2 + (P(2) - P(-2)))
                   + GAM * (C(l, 3) * unp3 - C(l, -3) * unm3 + (P(3) - P(-3)))
                   + DEL * (C(l, 4) * unp4 - C(l, -4) * unm4 + (P(4) - P(-4)));
        else
            A[l] = ALP * (C(l, 1) * unp1 - C(l, -1) * unm1)
                   + BET * (C(l, 2) * unp2 - C(l, -2) * unm2)
                   + GAM * (C(l, 3) * unp3 - C(l, -3) * unm3)
                   + DEL * (C(l, 4) * unp4 - C(l, -4) * unm4);
    }
    A[iene] = ALP * (C(iene, 1) * unp1 - C(iene, -1) * unm1 + (P(1) * unp1 - P(-1) * unm1))
              + BET * (C(iene, 2) * unp2 - C(iene, -2) * unm2 + (P(2) * unp2 - P(-2) * unm2))
              + GAM * (C(iene, 3) * unp3 - C(iene, -3) * unm3 + (P(3) * unp3 - P(-3) * unm3))
              + DEL * (C(iene, 4) * unp4 - C(iene, -4) * unm4 + (P(4) * unp4 - P(-4) * unm4));
#undef C
#undef QN
#undef P
}

/* cpu/hypterm.cpp:11-209 */
int orc_hypterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                const double *cons, const double *q, double *flux)
{
    layout_t L;
    double dxinv[3];
    int i, j, k, l, d;
    if (make_layout(lo, hi, ng, &L)) return -1;
    for (d = 0; d < 3; d++) dxinv[d] = 1.0E0 / dx[d];

#pragma omp parallel for collapse(2) private(i, l)
    for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (i = lo[0]; i <= hi[0]; i++) {
                long g = GIDX(&L, k, j, i);
                long o = IIDX(&L, k - ng, j - ng, i - ng);
                double A[5];
                hyp_sums(cons, q, g, 1, L.cg, qu, imx, A);
                for (l = 0; l < NC; l++) flux[l * L.c + o] = -(A[l]) * dxinv[0];
                hyp_sums(cons, q, g, L.xg, L.cg, qv, imy, A);
                for (l = 0; l < NC; l++) flux[l * L.c + o] = flux[l * L.c + o] - (A[l]) * dxinv[1];
                hyp_sums(cons, q, g, L.pg, L.cg, qw, imz, A);
                for (l = 0; l < NC; l++) flux[l * L.c + o] = flux[l * L.c + o] - (A[l]) * dxinv[2];
            }
    return 0;
}

/* -------------------------------------------------------------- diffterm */
/* 8th-order first derivative of f at g along stride s, before *dxinv
 * (cpu/diffterm.cpp:77-81 and every other first-derivative block). */
static inline double d1(const double *f, long g, long s)
{
    return ALP * (f[g + s] - f[g - s])
           + BET * (f[g + 2 * s] - f[g - 2 * s])
           + GAM * (f[g + 3 * s] - f[g - 3 * s])
           + DEL * (f[g + 4 * s] - f[g - 4 * s]);
}

/* 8th-order second derivative, before *SQR(dxinv) (cpu/diffterm.cpp:159-163). */
static inline double d2(const double *f, long g, long s)
{
    const double CENTER = -205.0E0 / 72.0E0;
    const double OFF1 = 8.0E0 / 5.0E0;
    const double OFF2 = -0.2E0;
    const double OFF3 = 8.0E0 / 315.0E0;
    const double OFF4 = -1.0E0 / 560.0E0;
    return CENTER * f[g]
           + OFF1 * (f[g + s] + f[g - s])
           + OFF2 * (f[g + 2 * s] + f[g - 2 * s])
           + OFF3 * (f[g + 3 * s] + f[g - 3 * s])
           + OFF4 * (f[g + 4 * s] + f[g - 4 * s]);
}

/* cpu/diffterm.cpp:22-320 */
int orc_diffterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                 const double *q, double *difflux, double eta, double alam)
{
    const double OneThird = 1.0E0 / 3.0E0;
    const double TwoThirds = 2.0E0 / 3.0E0;
    const double FourThirds = 4.0E0 / 3.0E0;
    layout_t L;
    double dxinv[3];
    double *tmp, *ux, *vx, *wx, *uy, *vy, *wy, *uz, *vz, *wz;
    const double *qU, *qV, *qW, *qT;
    long sx, sy, sz, c;
    int i, j, k, d;
    if (make_layout(lo, hi, ng, &L)) return -1;
    for (d = 0; d < 3; d++) dxinv[d] = 1.0E0 / dx[d];
    sx = 1;
    sy = L.xg;
    sz = L.pg;
    c = L.c;
    qU = q + qu * L.cg;
    qV = q + qv * L.cg;
    qW = q + qw * L.cg;
    qT = q + qtemp * L.cg;

    /* nine ghosted temporaries, as cpu/diffterm.cpp:59-61 */
    tmp = (double *)calloc((size_t)9 * L.cg, sizeof(double));
    if (!tmp) return -2;
    ux = tmp; vx = ux + L.cg; wx = vx + L.cg;
    uy = wx + L.cg; vy = uy + L.cg; wy = vy + L.cg;
    uz = wy + L.cg; vz = uz + L.cg; wz = vz + L.cg;

    /* cpu/diffterm.cpp:63-68 */
#pragma omp parallel for collapse(2) private(i)
    for (k = 0; k < L.n[2]; k++)
        for (j = 0; j < L.n[1]; j++)
            for (i = 0; i < L.n[0]; i++) difflux[IIDX(&L, k, j, i)] = 0.0E0;

    /* :73-97  x-derivatives on i interior, j and k ghosted */
#pragma omp parallel for collapse(2) private(i)
    for (k = lo[2] - ng; k <= hi[2] + ng; k++)
        for (j = lo[1] - ng; j <= hi[1] + ng; j++)
            for (i = lo[0]; i <= hi[0]; i++) {
                long g = GIDX(&L, k, j, i);
                ux[g] = d1(qU, g, sx) * dxinv[0];
                vx[g] = d1(qV, g, sx) * dxinv[0];
                wx[g] = d1(qW, g, sx) * dxinv[0];
            }
    /* :100-124  y-derivatives on j interior */
#pragma omp parallel for collapse(2) private(i)
    for (k = lo[2] - ng; k <= hi[2] + ng; k++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (i = lo[0] - ng; i <= hi[0] + ng; i++) {
                long g = GIDX(&L, k, j, i);
                uy[g] = d1(qU, g, sy) * dxinv[1];
                vy[g] = d1(qV, g, sy) * dxinv[1];
                wy[g] = d1(qW, g, sy) * dxinv[1];
            }
    /* :127-151  z-derivatives on k interior */
#pragma omp parallel for collapse(2) private(i)
    for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1] - ng; j <= hi[1] + ng; j++)
            for (i = lo[0] - ng; i <= hi[0] + ng; i++) {
                long g = GIDX(&L, k, j, i);
                uz[g] = d1(qU, g, sz) * dxinv[2];
                vz[g] = d1(qV, g, sz) * dxinv[2];
                wz[g] = d1(qW, g, sz) * dxinv[2];
            }

#pragma omp parallel for collapse(2) private(i)
    for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (i = lo[0]; i <= hi[0]; i++) {
                long g = GIDX(&L, k, j, i);
                long o = IIDX(&L, k - ng, j - ng, i - ng);
                double uxx, uyy, uzz, vyx, wzx, vxx, vyy, vzz, uxy, wzy, wxx, wyy, wzz, uxz, vyz;
                double txx, tyy, tzz, divu, tauxx, tauyy, tauzz, tauxy, tauxz, tauyz, mechwork;

                /* :159-187 */
                uxx = d2(qU, g, sx) * (dxinv[0] * dxinv[0]);
                uyy = d2(qU, g, sy) * (dxinv[1] * dxinv[1]);
                uzz = d2(qU, g, sz) * (dxinv[2] * dxinv[2]);
                vyx = d1(vy, g, sx) * dxinv[0];
                wzx = d1(wz, g, sx) * dxinv[0];
                difflux[o + imx * c] = eta * (FourThirds * uxx + uyy + uzz + OneThird * (vyx + wzx));

                /* :198-226 */
                vxx = d2(qV, g, sx) * (dxinv[0] * dxinv[0]);
                vyy = d2(qV, g, sy) * (dxinv[1] * dxinv[1]);
                vzz = d2(qV, g, sz) * (dxinv[2] * dxinv[2]);
                uxy = d1(ux, g, sy) * dxinv[1];
                wzy = d1(wz, g, sy) * dxinv[1];
                difflux[o + imy * c] = eta * (vxx + FourThirds * vyy + vzz + OneThird * (uxy + wzy));

                /* :237-265 */
                wxx = d2(qW, g, sx) * (dxinv[0] * dxinv[0]);
                wyy = d2(qW, g, sy) * (dxinv[1] * dxinv[1]);
                wzz = d2(qW, g, sz) * (dxinv[2] * dxinv[2]);
                uxz = d1(ux, g, sz) * dxinv[2];
                vyz = d1(vy, g, sz) * dxinv[2];
                difflux[o + imz * c] = eta * (wxx + wyy + FourThirds * wzz + OneThird * (uxz + vyz));

                /* :276-311 */
                txx = d2(qT, g, sx) * (dxinv[0] * dxinv[0]);
                tyy = d2(qT, g, sy) * (dxinv[1] * dxinv[1]);
                tzz = d2(qT, g, sz) * (dxinv[2] * dxinv[2]);
                divu = TwoThirds * (ux[g] + vy[g] + wz[g]);
                tauxx = 2.E0 * ux[g] - divu;
                tauyy = 2.E0 * vy[g] - divu;
                tauzz = 2.E0 * wz[g] - divu;
                tauxy = uy[g] + vx[g];
                tauxz = uz[g] + wx[g];
                tauyz = vz[g] + wy[g];
                mechwork = tauxx * ux[g] + tauyy * vy[g] + tauzz * wz[g]
                           + (tauxy * tauxy) + (tauxz * tauxz) + (tauyz * tauyz);
                mechwork = eta * mechwork
                           + difflux[o + imx * c] * qU[g]
                           + difflux[o + imy * c] * qV[g]
                           + difflux[o + imz * c] * qW[g];
                difflux[o + iene * c] = alam * (txx + tyy + tzz) + mechwork;
            }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------ RK update */
/* cpu/advance.cpp:78-85, 110-119, 144-153 */
int orc_rk_update(int phase, const int lo[3], const int hi[3], int ng,
                  double *u, double *unew, const double *d, const double *f, double dt)
{
    const double OneThird = 1.0E0 / 3.0E0;
    const double TwoThirds = 2.0E0 / 3.0E0;
    const double OneQuarter = 1.E0 / 4.E0;
    const double ThreeQuarters = 3.E0 / 4.E0;
    layout_t L;
    int i, j, k, l;
    if (make_layout(lo, hi, ng, &L)) return -1;
    if (phase < 1 || phase > 3) return -1;
#pragma omp parallel for collapse(3) private(i)
    for (l = 0; l < NC; l++)
        for (k = 0; k < L.n[2]; k++)
            for (j = 0; j < L.n[1]; j++)
                for (i = 0; i < L.n[0]; i++) {
                    long g = l * L.cg + GIDX(&L, k + ng, j + ng, i + ng);
                    long o = l * L.c + IIDX(&L, k, j, i);
                    if (phase == 1)
                        unew[g] = u[g] + dt * (d[o] + f[o]);
                    else if (phase == 2)
                        unew[g] = ThreeQuarters * u[g] + OneQuarter * (unew[g] + dt * (d[o] + f[o]));
                    else
                        u[g] = OneThird * u[g] + TwoThirds * (unew[g] + dt * (d[o] + f[o]));
                }
    return 0;
}

/* --------------------------------------------------------------- advance */
/* cpu/advance.cpp:13-154 (new_advance) */
int orc_advance(int n, const double dx[3], double cfl, double eta, double alam,
                double *U, double *Unew, double *Q, double *D, double *F, double *dt)
{
    int lo[3] = {NG, NG, NG}, hi[3], dim[3], dim_g[3], d, rc = 0;
    double courno;
    for (d = 0; d < 3; d++) {
        hi[d] = n - 1 + NG;
        dim[d] = n;
        dim_g[d] = n + 2 * NG;
    }
    /* stage 1 */
    rc |= orc_fill_boundary(U, dim, dim_g);
    courno = 1.0E-50;
    rc |= orc_ctoprim(lo, hi, U, Q, dx, NG, &courno);
    *dt = cfl / courno;
    rc |= orc_diffterm(lo, hi, NG, dx, Q, D, eta, alam);
    rc |= orc_hypterm(lo, hi, NG, dx, U, Q, F);
    rc |= orc_rk_update(1, lo, hi, NG, U, Unew, D, F, *dt);
    /* stage 2 */
    rc |= orc_fill_boundary(Unew, dim, dim_g);
    rc |= orc_ctoprim(lo, hi, Unew, Q, dx, NG, NULL);
    rc |= orc_diffterm(lo, hi, NG, dx, Q, D, eta, alam);
    rc |= orc_hypterm(lo, hi, NG, dx, Unew, Q, F);
    rc |= orc_rk_update(2, lo, hi, NG, U, Unew, D, F, *dt);
    /* stage 3 */
    rc |= orc_fill_boundary(Unew, dim, dim_g);
    rc |= orc_ctoprim(lo, hi, Unew, Q, dx, NG, NULL);
    rc |= orc_diffterm(lo, hi, NG, dx, Q, D, eta, alam);
    rc |= orc_hypterm(lo, hi, NG, dx, Unew, Q, F);
    rc |= orc_rk_update(3, lo, hi, NG, U, Unew, D, F, *dt);
    return rc ? -1 : 0;
}

/* ------------------------------------------------------------- threading */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
