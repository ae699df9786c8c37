// cns_batch_dev.h -- device records and launchers of the batched engine (internal).
//
// Boxes live in one device pool, ghost-free: component l of a box of edge e is
// e^3 doubles, index (k*e + j)*e + i.  The periodic wrap is done by index when
// halos are gathered, so no ghost storage and no fill pass exist.
#pragma once

#include "cns_common.h"

namespace cns {

struct BoxDev {
    int e;          // edge (cells)
    int err;        // 1: non-finite Courant number seen
    double dxi[3];  // 1/dx
    double dx[3];
    double cfl, eta, alam;
    double pref;    // pressure reference subtracted before the flux arithmetic (0 in FP64 mode)
    double *U;      // U^n, 5 comps
    double *S1;     // U^{1/3}
    double *S2;     // U^{2/3}
    void *Q[3];     // primitives (u, v, w, p - pref, T) of U^n, U^{1/3}, U^{2/3}
    void *D;        // diffterm of the current stage (mx, my, mz, E), flux type
    void *WZ;       // wz = d(w)/dz of the current stage's primitives, flux type
    unsigned long long acc;  // Courant maximum (bits) of the state the next step starts from
    double dt;      // dt of the current step
    double courno;  // its Courant number
    long long steps;
    long long target;  // the box stops advancing once steps reaches it (0: no limit)
};

// A box that has reached its step target is skipped by every kernel of the step.
__host__ __device__ inline bool box_stopped(const BoxDev &b) { return b.target > 0 && b.steps >= b.target; }

struct TileDev {
    int box;
    short x0, y0, z0, nz;  // tile origin (cells) and z run
    short ltx;             // log2 of the tile width (4, 5 or 6); height = 256 >> ltx
    short pad;
};

// Tile widths 16, 32, 64 (class 0, 1, 2); a box's tiles all have one width.
constexpr int kTileClasses = 3;

constexpr int kBatchThreads = 256;

hipError_t launch_batch_init(BoxDev *boxes, int box, int e, bool fp32, hipStream_t s);
hipError_t launch_batch_finalize(BoxDev *boxes, const int *box_ids, int nbox, int count_step,
                                 hipStream_t s);
// One RK stage over the tile list: the list holds class 0, 1, 2 tiles in the
// ranges [start[c], start[c]+count[c]); each class gets a diffterm launch, then a
// hypterm + update launch, of one workgroup per tile.  staged_diff selects the
// diffterm whose planes are staged by direct-to-LDS loads, preceded by a wz-field
// launch per class; keep_centre the FP32 hypterm + update of stages 2 and 3 that keeps
// its window's FP64 leading values for the centre plane; staged_hyp (FP32 flux) and
// staged_hyp64 (FP64) the hypterm + update of 32-wide tiles whose planes are staged by
// direct-to-LDS loads.
hipError_t launch_batch_stage(BoxDev *boxes, const TileDev *tiles, const int start[3],
                              const int count[3], int stage, bool fp32, bool staged_diff,
                              bool keep_centre, bool staged_hyp, bool staged_hyp64, hipStream_t s);

}  // namespace cns
