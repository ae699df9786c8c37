/*
 * cns_domain.h -- C ABI of multi-box domains in libcns_hip.so: one periodic domain
 * cut into boxes, each box a ghosted array of the reference layout, advanced
 * together with one dt.
 *
 * Replaces the reference's BoxLib domain setup and step (fortran90/main.f90:97-116:
 * make_box of n_cell cells, boxarray_build_bx, boxarray_maxsize(max_grid_size),
 * layout_build_ba with periodic pmask, multifab_build(U, la, NC, NG);
 * fortran90/advance.f90:81-139: multifab_fill_boundary, then per box ctoprim with
 * the Courant number reduced over every box by parallel_reduce(MPI_MAX), dt =
 * cfl/courno, diffterm, hypterm and the SSP-RK3 updates per box).  Every box takes
 * the same per-cell arithmetic as cns_advance, and its ghost cells are filled
 * from the neighbouring boxes, so a domain's state after any number of steps is
 * bit-identical to a single box of the whole domain advanced by cns_advance.
 *
 * Boxes: each extent ncell[d] is cut into ncell[d] / max_grid_size pieces of
 * max_grid_size cells (ncell[d] must be a multiple of max_grid_size, which must
 * be >= 4).  Box n = I + nb[0]*(J + nb[1]*K) covers cells I*m .. I*m+m-1 along x
 * (m = max_grid_size), likewise J along y and K along z.
 *
 * Arrays: a domain array stores its boxes back to back, box n at offset
 * n * cns_domain_array_len(d, ncomp, ghosted) / nbox doubles, each box in the
 * layout of include/cns.h for an m x m x m box (ghosted: U, Unew with 5
 * components, Q with 6; interior: D and F with 5).  Every array must be 16-byte
 * aligned (the pipelined kernels move 16 bytes per lane).
 *
 * All calls use the calling thread's current stream (cns_set_stream).  Calls on
 * one domain must not overlap.
 */
#ifndef CNS_DOMAIN_H
#define CNS_DOMAIN_H

#include <stddef.h>

#include "cns.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cns_domain cns_domain;

/* A periodic domain of ncell[0] x ncell[1] x ncell[2] cells of size dx, cut into
 * boxes of max_grid_size^3 cells; cfl, eta, alam as cns_ctx_create.  Allocates the
 * domain's scratch (a stage buffer and the wz field, 6 ghosted components per box). */
int cns_domain_create(cns_domain **d, const int ncell[3], int max_grid_size, const double dx[3],
                      double cfl, double eta, double alam);
int cns_domain_destroy(cns_domain *d);

/* Number of boxes along each axis (nb) and in all (*nbox). */
int cns_domain_boxes(const cns_domain *d, int nb[3], int *nbox);

/* Cell range of box n, 0-based and inclusive (BoxLib's lwb/upb of get_box). */
int cns_domain_box(const cns_domain *d, int n, int lo[3], int hi[3]);

/* Doubles of a domain array of ncomp components, ghosted (1) or interior (0). */
size_t cns_domain_array_len(const cns_domain *d, int ncomp, int ghosted);

/* multifab_fill_boundary (fortran90/advance.f90:81) of a 5-component ghosted domain
 * array: every ghost cell of every box gets the domain cell at the periodic image
 * of its index, from the box that holds it. */
int cns_domain_fill_boundary(cns_domain *d, double *U);

/* One SSP-RK3 step of every box (fortran90/advance.f90:57-410; per box the same
 * composition as cns_advance), with dt = cfl / (the Courant maximum over every
 * box's interior).  U in/out; Unew, Q, D, F are the caller's domain arrays and end
 * as cns_advance leaves them, box by box.  Writes dt to *dt (host) and synchronises
 * the stream; CNS_ENUMERIC for a non-finite Courant number. */
int cns_domain_advance(cns_domain *d, double *U, double *Unew, double *Q, double *D, double *F,
                       double *dt);

#ifdef __cplusplus
}
#endif

#endif /* CNS_DOMAIN_H */
