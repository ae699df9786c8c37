/*
 * cns_oracle.h -- CPU restatement of the CNS_NoSpec time step (TEST INFRASTRUCTURE).
 *
 * This is the parity oracle: a scalar, single-threaded C restatement of the
 * reference's cpu/ path, written in the padded SoA ("ghosted") layout that the
 * HIP library's C ABI uses.  Only tests, the smoke check and the CPU-baseline
 * leg of the benchmark may call it.  The product path never does.
 *
 * Every function follows the operation order of the reference file:line
 * cited beside it, so that, built with -ffp-contract=off, it reproduces the
 * reference bit for bit.  Layout (gpu/header.h:14, gpu/util_cpu.cpp:327-343):
 *   ghosted  : idx = l*cg + k*pg + j*xg + i, xg = PAD16(n0+2ng), pg = xg*(n1+2ng), cg = pg*(n2+2ng)
 *   interior : idx = l*c  + k*p  + j*x  + i, x  = PAD16(n0),     p  = x*n1,        c  = p*n2
 * Index space is ghost-shifted as in cpu/advance.cpp:33-34: lo = ng, hi = n-1+ng.
 */
#ifndef CNS_ORACLE_H
#define CNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* cpu/helper_functions.cpp:226-262 (3 ordered passes z, y, x).  dim = interior, dim_g = dim + 8. */
int orc_fill_boundary(double *u, const int dim[3], const int dim_g[3]);

/* cpu/ctoprim.cpp:11-105.  courno == NULL selects the overload without the Courant reduction. */
int orc_ctoprim(const int lo[3], const int hi[3], const double *u, double *q,
                const double dx[3], int ng, double *courno);

/* cpu/hypterm.cpp:11-209 */
int orc_hypterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                const double *cons, const double *q, double *flux);

/* cpu/diffterm.cpp:22-320 */
int orc_diffterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                 const double *q, double *difflux, double eta, double alam);

/* cpu/advance.cpp:78-85 (phase 1), 110-119 (phase 2), 144-153 (phase 3).
 * phase 1: unew = u + dt(d+f); phase 2: unew = 3/4 u + 1/4(unew + dt(d+f));
 * phase 3: u = 1/3 u + 2/3(unew + dt(d+f)).  Interior cells only. */
int orc_rk_update(int phase, const int lo[3], const int hi[3], int ng,
                  double *u, double *unew, const double *d, const double *f, double dt);

/* cpu/advance.cpp:13-154 (new_advance).  One RK3 step; dt is an output. */
int orc_advance(int n, const double dx[3], double cfl, double eta, double alam,
                double *U, double *Unew, double *Q, double *D, double *F, double *dt);

/* Threads of the per-cell loops (n > 0 sets it); returns the count in use.  Results
 * do not depend on it. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
