// cns_kernels.h -- host launchers of the CNS_NoSpec HIP kernels (internal).
#pragma once

#include "cns_common.h"

namespace cns {

// A set of n boxes of one geometry stored back to back: box b of an array starts
// b * stride doubles after box 0, with stride su for 5-component ghosted arrays (U),
// sq for 6-component ghosted ones (Q), sc for 5-component interior ones (D, F) and
// sw for the one-component ghosted wz field.  The default is a single box.
struct BoxSet {
    int n = 1;
    int64_t su = 0, sq = 0, sc = 0, sw = 0;
};
inline BoxSet box_set(const Geom &g, int n)
{
    BoxSet b;
    b.n = n;
    b.su = NC * g.cg;
    b.sq = NQ * g.cg;
    b.sc = NC * g.c;
    b.sw = g.cg;
    return b;
}

// Ghost shell, closed-form periodic image (cns_kernels.hip).
hipError_t launch_fill_boundary(const Geom &g, double *u, hipStream_t s);
// The same periodic fill of u's NC components and of the components of q selected by
// qmask (bit l for component l), in one pass.
hipError_t launch_fill_boundary_q(const Geom &g, double *u, double *q, unsigned qmask, hipStream_t s);
// The periodic fill of a domain of nb[0] x nb[1] x nb[2] boxes (box b = I + nb[0]*(J +
// nb[1]*K) at cell origin (I*n0, J*n1, K*n2)): each ghost cell of each box gets the
// domain cell at the periodic image of its global index, from whichever box holds
// it (multifab_fill_boundary, fortran90/advance.f90:81); u and the qmask components
// of q, as launch_fill_boundary_q.  With one box it equals launch_fill_boundary_q.
hipError_t launch_domain_fill(const Geom &g, const BoxSet &bs, const int nb[3], double *u, double *q,
                              unsigned qmask, hipStream_t s);

// Resets a Courant accumulator slot to the bits of max(init, 0) (cns_kernels.hip).
hipError_t launch_courno_reset(unsigned long long *acc, double init, hipStream_t s);

// ctoprim over the ghosted box; acc != nullptr adds the Courant max reduction.
// full = false stores only u, v, w and T (what diffterm reads), not rho and p.
hipError_t launch_ctoprim(const Geom &g, const double *u, double *q, unsigned long long *acc,
                          double dx0, double dx1, double dx2, hipStream_t s, bool full = true);

// dt = cfl / max(acc, 1e-50) into dt_dev (and the Courant number into courno_dev).
hipError_t launch_courno_to_dt(const unsigned long long *acc, double cfl, double *dt_dev,
                               double *courno_dev, hipStream_t s);

// RK3 stage update; dt from dt_dev when non-null, else dt_host.
hipError_t launch_rk_update(const Geom &g, int phase, double *u, double *unew, const double *d,
                            const double *f, const double *dt_dev, double dt_host, hipStream_t s);

// Reference-order restatements, one thread per output cell, direct global loads
// (cns_kernels.hip).  Used by the free-standing ABI entry points' cross-check path.
hipError_t launch_hypterm_direct(const Geom &g, const double *cons, const double *q, double *flux,
                                 double dxi0, double dxi1, double dxi2, hipStream_t s);
hipError_t launch_diffterm_direct(const Geom &g, const double *q, double *difflux, double dxi0,
                                  double dxi1, double dxi2, double eta, double alam, hipStream_t s);

// z-marching, LDS-tiled stencil kernels (cns_stencil.hip).
hipError_t launch_hypterm_tiled(const Geom &g, const double *cons, const double *q, double *flux,
                                double dxi0, double dxi1, double dxi2, hipStream_t s);
// hypterm with the RK3 stage update of its cell fused in: flux as cns_hypterm, then
// phase 1: out = cons + dt(d+f); 2: out = 3/4 base + 1/4(cons + dt(d+f));
// 3: out = 1/3 base + 2/3(cons + dt(d+f)).  out must not alias cons.
hipError_t launch_hypterm_rk(const Geom &g, const double *cons, const double *q, double *flux,
                             double dxi0, double dxi1, double dxi2, const double *d,
                             const double *base, double *out, const double *dt_dev, int phase,
                             hipStream_t s);
hipError_t launch_diffterm_tiled(const Geom &g, const double *q, double *difflux, double dxi0,
                                 double dxi1, double dxi2, double eta, double alam, hipStream_t s);

// DMA-pipelined z-marching stencils (cns_stencil_pipe.hip).  Every array must be
// 16-byte aligned (pipe_aligned).  hypterm: phase 0 is plain hypterm, 1..3 fuse the
// RK stage update as launch_hypterm_rk (phase 1 with base == cons takes the direct
// form); write_f = false skips storing flux (for a stage whose F is overwritten
// later).  diffterm reads wz, a one-component ghosted array filled by launch_wz from
// the same q.
bool pipe_aligned(const void *p);


hipError_t launch_hypterm_pipe(const Geom &g, const double *cons, const double *q, double *flux,
                               double dxi0, double dxi1, double dxi2, const double *d,
                               const double *base, double *out, const double *dt_dev, int phase,
                               bool write_f, hipStream_t s);
// The same fused stage reading only cons (phases 1..3): u, v, w, p are derived from it
// with ctoprim's arithmetic, so it equals launch_hypterm_pipe with q = ctoprim(cons).
// qout (with write_f false): ctoprim of the result on interior cells into qout, u, v,
// w and T (qout_kind 1) or all six components (2); 0 or a null qout stores none.
hipError_t launch_hypterm_cons(const Geom &g, const double *cons, double *flux, double dxi0,
                               double dxi1, double dxi2, const double *d, const double *base,
                               double *out, const double *dt_dev, int phase, bool write_f,
                               double *qout, int qout_kind, hipStream_t s, const BoxSet &bs = BoxSet());
hipError_t launch_wz(const Geom &g, const double *q, double *wz, double dxi2, hipStream_t s,
                     const BoxSet &bs = BoxSet());
// ctoprim (as launch_ctoprim, full or u, v, w, T only) that also writes the wz field of
// launch_wz from the same state, in one z-marching pass.
hipError_t launch_ctoprim_wz(const Geom &g, const double *u, double *q, double *wz,
                             unsigned long long *acc, double dx0, double dx1, double dx2, double dxi2,
                             bool full, hipStream_t s, const BoxSet &bs = BoxSet());
// write_rho = false leaves difflux's rho row (always 0.0) unwritten, for a caller that
// stores it once and never reads it in between.
hipError_t launch_diffterm_pipe(const Geom &g, const double *q, const double *wz, double *difflux,
                                double dxi0, double dxi1, double dxi2, double eta, double alam,
                                hipStream_t s, bool write_rho = true, const BoxSet &bs = BoxSet());

}  // namespace cns
