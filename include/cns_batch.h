/*
 * cns_batch.h -- C ABI of the batched engine of libcns_hip.so: many independent
 * periodic boxes of uneven edges resident in HBM and advanced together.
 *
 * No counterpart in the reference, which advances one box per process
 * (cpu/main.cpp:28-29, gpu/main.cu:48-49); each box here is the reference's
 * problem (cpu/advance.cpp:13-154: fill, ctoprim + Courant dt, diffterm,
 * hypterm, SSP-RK3), with its own dt from its own Courant number every step.
 *
 * Box state is ghost-free: a box of edge e holds its 5 conserved components as
 * e^3 doubles each, index (k*e + j)*e + i, component outermost (the periodic
 * wrap is done by index).  A step runs, per RK stage, a diffterm launch and a
 * hypterm + RK-update launch over a tile work list of every active box (one
 * launch per tile-width class, a workgroup per tile), then a dt finalize; it is captured
 * once into a hipGraph and replayed.  State, RK3 accumulation, p/T recovery and
 * the dt reduction are always FP64; with flux_fp32 the stencil arithmetic of
 * hypterm and diffterm runs in FP32 (the FP64 engine is bit-identical to
 * cns_advance).  Memory comes from one pool of hbm_budget bytes reserved at
 * creation; a box that does not fit is refused with CNS_EBUDGET.
 *
 * All calls use the calling thread's current stream (cns_set_stream).  Calls
 * on one engine must not overlap.
 */
#ifndef CNS_BATCH_H
#define CNS_BATCH_H

#include <stddef.h>

#include "cns.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CNS_EBUDGET (-5) /* the box does not fit in the engine's HBM budget */

typedef struct cns_batch cns_batch;

/* Reserve hbm_budget bytes of device memory for box state; max_boxes bounds the
 * number of boxes resident at once. */
int cns_batch_create(cns_batch **b, size_t hbm_budget, int flux_fp32, int max_boxes);
int cns_batch_destroy(cns_batch *b);

/* Device bytes one box of edge e takes in an engine of this precision. */
size_t cns_batch_box_bytes(int edge, int flux_fp32);

/* Admit a box: copy its initial state u0 (DEVICE pointer, ghost-free layout
 * above, 5*e^3 doubles), recover its primitives and its first dt.  dx is the cell
 * size (cube cells), cfl/eta/alam as the reference's general_input.  Returns the
 * box id in *box_id, or CNS_EBUDGET if the pool cannot hold it.  Edge 8..512. */
int cns_batch_add(cns_batch *b, int edge, double dx, double cfl, double eta, double alam,
                  const double *u0, int *box_id);

/* Release a box's memory (its id may be reused). */
int cns_batch_remove(cns_batch *b, int box_id);

/* Copy a box's current U (ghost-free layout) to u (DEVICE pointer, 5*e^3 doubles). */
int cns_batch_read(cns_batch *b, int box_id, double *u);

/* Advance every resident box by nsteps RK3 steps (graph replays, asynchronous). */
int cns_batch_step(cns_batch *b, int nsteps);

/* Step target of a box: once its step count (steps since admission) reaches
 * target, every later cns_batch_step leaves the box untouched (its state, dt
 * and count stay as they are), so a server can replay many steps without
 * stopping at each box's end.  target = 0 removes the limit.  No counterpart in
 * the reference. */
int cns_batch_set_target(cns_batch *b, int box_id, long long target);

/* Per-box status for n ids (synchronises the stream): steps taken since
 * admission, dt of the last step, its Courant number, and err = 1 when a
 * non-finite Courant number was seen (the box's state is not physical; the
 * other boxes are unaffected).  Any output pointer may be NULL. */
int cns_batch_query(cns_batch *b, int n, const int *ids, long long *steps, double *dt,
                    double *courno, int *err);

/* Tile shape used for boxes of a given edge: width 16, 32 or 64 cells (height
 * 256/width) and z run kz (1..edge).  Applies to boxes admitted afterwards and
 * rebuilds the work list.  The table starts from built-in defaults; the
 * package's autotuner measures and sets it. */
int cns_batch_set_tile(cns_batch *b, int edge, int tile_width, int kz);
int cns_batch_get_tile(cns_batch *b, int edge, int *tile_width, int *kz);

/* Pool bytes in use, resident boxes, tiles in the work list. */
int cns_batch_stats(cns_batch *b, size_t *used_bytes, int *boxes, int *tiles);

#ifdef __cplusplus
}
#endif
#endif /* CNS_BATCH_H */
