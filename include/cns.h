/*
 * cns.h -- C ABI of the MI355X-native CNS_NoSpec time step (libcns_hip.so).
 *
 * The entry points keep the argument order and meaning of the reference's
 * cpu/ functions (cpu/header.h:97-156, cpu/helper_functions.h:15-19) with flat
 * pointers in place of `double ****` and `double *` in place of `double &`.
 * Arrays are DEVICE pointers in the reference GPU path's padded SoA layout
 * (gpu/header.h:14, gpu/util_cpu.cpp:327-343):
 *
 *   ghosted  U, Unew (5 comps), Q (6 comps): idx = l*cg + k*pg + j*xg + i
 *            xg = PAD16(n0+2ng), pg = xg*(n1+2ng), cg = pg*(n2+2ng)
 *   interior D, F (5 comps):                 idx = l*c + k*p + j*x + i
 *            x = PAD16(n0), p = x*n1, c = p*n2
 *
 * PAD16(v) rounds v up to a multiple of 16.  Index space is ghost-shifted:
 * lo[d] = ng, hi[d] = n[d]-1+ng.  ng must be 4 (the reference's NG) and every
 * extent n[d] >= 4.  Component order: U = (rho, rho u, rho v, rho w, rho E),
 * Q = (rho, u, v, w, p, T).
 *
 * Every call enqueues its kernels on the calling thread's current stream
 * (cns_set_stream; the null stream by default) and returns 0 or a negative
 * CNS_E* code; nothing calls exit().  Calls that hand a value back to the host
 * (a non-NULL courno, cns_advance's dt) synchronise that stream.
 */
#ifndef CNS_H
#define CNS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CNS_OK 0
#define CNS_EINVAL (-1)   /* bad argument or unsupported geometry */
#define CNS_EHIP (-2)     /* a HIP runtime call failed */
#define CNS_ENOMEM (-3)   /* device allocation failed */
#define CNS_ENUMERIC (-4) /* non-finite Courant number: state is not physical */

/* Replaces CUDA_SAFE_CALL's print-and-exit (gpu/util.cu:7-15): text of the
 * calling thread's last error. */
const char *cns_last_error_string(void);

/* Stream of the calling thread's later calls (a hipStream_t; NULL = null stream).
 * The reference always uses the default stream (gpu/advance.cu:107-203). */
int cns_set_stream(void *hip_stream);
void *cns_get_stream(void);

/* Device workspace for the free-standing per-kernel calls of the calling thread:
 * cns_ctoprim with a non-NULL courno needs one 8-byte slot (its Courant accumulator),
 * cns_diffterm one ghosted component of doubles (the wz field of the tiled kernel);
 * cns_scratch_bytes() gives the size for a box (0 for bad bounds).  With a workspace of
 * at least that size set on the current device (dev_ptr 256-byte aligned, owned and
 * freed by the caller, not used by it while the calls that read it run), those calls
 * allocate nothing.  Without one they allocate a per-thread, per-device buffer on first
 * use (grown when a larger box needs it) and keep it until the thread exits.
 * cns_set_scratch(NULL, 0) removes it.  cns_advance and the other context calls use the
 * context's own buffers.  No counterpart in the reference, whose per-call temporaries
 * are allocated by the caller at start-up (gpu/util.cu:171-175). */
int cns_set_scratch(void *dev_ptr, size_t bytes);
size_t cns_scratch_bytes(const int lo[3], const int hi[3], int ng);

/* Which kernels cns_hypterm, cns_diffterm and the step use (process-wide):
 * CNS_IMPL_TILED (default) -- z-marching LDS-tiled kernels;
 * CNS_IMPL_DIRECT -- one thread per cell with direct loads, the reference's
 * loops restated, kept as an independent GPU cross-check of the tiled kernels
 * at sizes the CPU oracle cannot reach.  Both give bit-identical results.
 * No counterpart in the reference. */
#define CNS_IMPL_TILED 0
#define CNS_IMPL_DIRECT 1
int cns_set_stencil_impl(int impl);
int cns_get_stencil_impl(void);

/* Periodic ghost fill, in place, over the 5 components of a ghosted array.
 * dim = interior extents, dim_g = dim + 8.  Bit-exact copy semantics of the
 * three ordered passes of the reference.
 * Replaces: fill_boundary(U, dim, dim_ng)  cpu/helper_functions.h:15-19 (body :226-262);
 *           gpu_fill_boundary(h, d, u)     gpu/util.h:17-21 (body gpu/util.cu:111-127). */
int cns_fill_boundary(double *u, const int dim[3], const int dim_g[3]);

/* Primitive recovery over the whole ghosted box, u (5) -> q (6).  When courno is
 * non-NULL (a HOST pointer) it is in/out: *courno = max(*courno, max over interior
 * cells of (c+|u_d|)/dx_d), and the stream is synchronised.  NULL selects the
 * reference overload without the Courant reduction.
 * Replaces: ctoprim(lo,hi,u,q,dx,ng,courno&)  cpu/header.h:99-107 (body cpu/ctoprim.cpp:11-70);
 *           ctoprim(lo,hi,u,q,dx,ng)          cpu/header.h:108-115 (body cpu/ctoprim.cpp:72-105);
 *           gpu_ctoprim(h,d,u,q[,courno&])    gpu/header.h:146-158. */
int cns_ctoprim(const int lo[3], const int hi[3], const double *u, double *q,
                const double dx[3], int ng, double *courno);

/* Inviscid flux divergence: cons (5, ghosted) and q (6, ghosted) -> flux (5, interior).
 * Replaces: hypterm(lo,hi,ng,dx,cons,q,flux)  cpu/header.h:118-126 (body cpu/hypterm.cpp:11-209);
 *           gpu_hypterm3(h,d,cons,q,flux)     gpu/header.h:192-198. */
int cns_hypterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                const double *cons, const double *q, double *flux);

/* Viscous and heat-conduction terms: q (6, ghosted) -> difflux (5, interior).
 * Replaces: diffterm(lo,hi,ng,dx,q,difflux,eta,alam)  cpu/header.h:129-138 (body cpu/diffterm.cpp:22-320);
 *           gpu_diffterm2(h,d,q,difflux,flux)         gpu/header.h:223-229. */
int cns_diffterm(const int lo[3], const int hi[3], int ng, const double dx[3],
                 const double *q, double *difflux, double eta, double alam);

/* cns_diffterm for callers that cannot allocate device memory: code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same arguments
 * and the same bits as cns_diffterm, but its scratch (the wz field of the tiled
 * kernel, one ghosted component) is a static device array of
 * CNS_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) that each GPU holds from the
 * library's first use until the process ends; nothing is allocated or freed.
 * Boxes whose ghosted component exceeds it (PAD16(n0+8)*(n1+8)*(n2+8) values;
 * cubes above 192^3) are rejected with CNS_EINVAL.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other threads, other streams) race
 * and must be serialised by the caller.
 * No counterpart in the reference. */
#define CNS_NOALLOC_SCRATCH_DOUBLES (8 << 20)
int cns_diffterm_noalloc(const int lo[3], const int hi[3], int ng, const double dx[3],
                         const double *q, double *difflux, double eta, double alam);

/* One SSP-RK3 stage update over interior cells (ghosts untouched):
 *   phase 1: unew = u + dt(d+f)
 *   phase 2: unew = 3/4 u + 1/4 (unew + dt(d+f))
 *   phase 3: u    = 1/3 u + 2/3 (unew + dt(d+f))
 * Replaces: gpu_Unew(h,d,unew,u,d,f,dt,phase)  gpu/header.h:319-328 (body gpu/advance.cu:13-105);
 *           the update loops of cpu/advance.cpp:78-85,110-119,144-153. */
int cns_rk_update(int phase, const int lo[3], const int hi[3], int ng,
                  double *u, double *unew, const double *d, const double *f, double dt);

/* Per-box configuration and device scratch: the reference's global h_const.
 * ncells = interior extents, dx = cell sizes, cfl/eta/alam as in the reference's
 * general_input.  Allocates a small device block and the stage buffer of the
 * fused step (see cns_ctx_set_fused); no allocation happens on the step path.
 * Replaces: read_configurations + h_const  cpu/helper_functions.cpp:38-83, cpu/main.cpp:4;
 *           copy_configurations + d_const  gpu/util.cu:129-147, gpu/main.cu:6-9. */
typedef struct cns_ctx cns_ctx;
int cns_ctx_create(cns_ctx **ctx, const int ncells[3], const double dx[3],
                   double cfl, double eta, double alam);
int cns_ctx_destroy(cns_ctx *ctx);

/* Selects how cns_advance composes a stage.  fused = 1 (the default set by
 * cns_ctx_create): fill, ctoprim, diffterm, then hypterm with the RK update of each
 * cell applied in the same pass; the context then owns one ghosted 5-component
 * stage buffer (the U^{1/3} state), allocated here or at creation.  fused = 0:
 * kernel for kernel as the reference (separate hypterm and update launches), and
 * the stage buffer is released.  Both give bit-identical results.  No
 * counterpart in the reference. */
int cns_ctx_set_fused(cns_ctx *ctx, int fused);

/* One full RK3 time step: fill_boundary, ctoprim (+courno), dt = cfl/courno,
 * diffterm, hypterm, update, three times, dt fixed after stage 1.  On return U
 * holds U^{n+1} (ghosts: images of U^n, as the reference leaves them), Unew holds
 * U^{2/3} with ghosts filled, Q/D/F hold the stage-3 values, and *dt (HOST
 * pointer) the step's dt.  Synchronises the stream.  Returns CNS_ENUMERIC if the
 * Courant number is not finite.
 * Replaces: new_advance(U,Unew,Q,D,F,dt&)        cpu/header.h:149-156 (body cpu/advance.cpp:13-154);
 *           gpu_advance(h,d,U,Unew,Q,D,F,dt&)    gpu/header.h:287-296 (body gpu/advance.cu:107-203). */
int cns_advance(cns_ctx *ctx, double *U, double *Unew, double *Q, double *D, double *F,
                double *dt);

/* The same step with dt left in DEVICE memory (*dt_dev, one double) and no
 * synchronisation, so that a caller can capture steps in a hipGraph.  The
 * Courant number of the step's first stage is left in *courno_dev (one double,
 * device) when courno_dev is non-NULL.  No counterpart in the reference, whose
 * step always synchronises on the host (gpu/advance.cu:130-135). */
int cns_advance_async(cns_ctx *ctx, double *U, double *Unew, double *Q, double *D, double *F,
                      double *dt_dev, double *courno_dev);

/* Live per-kernel timing of the context's steps: when enabled, each kernel launch
 * of cns_advance/cns_advance_async is bracketed by HIP events on its stream.
 * cns_ctx_profile() turns timing on or off and clears the totals;
 * cns_ctx_profile_read() waits for the recorded launches and returns, per kernel
 * kind, the summed milliseconds and the number of launches.  Must not be on
 * while a step is being captured into a graph.  No counterpart in the
 * reference (whose timers are host gettimeofday, cpu/advance.cpp:53-73). */
#define CNS_PROF_FILL 0
#define CNS_PROF_CTOPRIM 1
#define CNS_PROF_DIFFTERM 2
#define CNS_PROF_HYPTERM 3
#define CNS_PROF_RK 4
#define CNS_PROF_HYPTERM_RK 5
#define CNS_PROF_FILL_CTOPRIM 6
#define CNS_PROF_OTHER 7
#define CNS_PROF_NKINDS 8
int cns_ctx_profile(cns_ctx *ctx, int enable);
int cns_ctx_profile_read(cns_ctx *ctx, double ms_total[CNS_PROF_NKINDS],
                         long long launches[CNS_PROF_NKINDS]);

#ifdef __cplusplus
}
#endif
#endif /* CNS_H */
