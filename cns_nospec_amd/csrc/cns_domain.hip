// cns_domain.hip -- multi-box periodic domains (include/cns_domain.h).
//
// A domain's boxes all have one geometry and sit back to back in each array, so
// every kernel of the step covers every box in one launch: the pipelined stencil
// kernels take the box as the leading part of their workgroup index, the column
// kernels (ctoprim + wz, wz) as their third grid dimension, and the fill copies
// ghost cells across boxes.  The step is the fused composition of cns_advance
// (cns_abi.hip) with the box fill in place of the single-box fill; the Courant
// maxima of all boxes meet in one accumulator, as BoxLib's parallel_reduce does
// (fortran90/advance.f90:139).
#include <cmath>
#include <new>

#include "cns_domain.h"
#include "cns_host.h"
#include "cns_kernels.h"

using namespace cns;

struct cns_domain {
    Geom g;          // one box
    BoxSet bs;       // strides of the box-stacked arrays
    int nb[3];       // boxes per axis
    int m;           // max_grid_size
    double dx[3], dxi[3];
    double cfl, eta, alam;
    unsigned long long *d_acc;  // [0] Courant bits, [1] dt, [2] courno
    double *d_dt, *d_courno;
    double *d_stage;  // U^{1/3} of every box (5 ghosted comps per box)
    double *d_wz;     // wz field of every box
};

extern "C" {

int cns_domain_create(cns_domain **out, const int ncell[3], int max_grid_size, const double dx[3],
                      double cfl, double eta, double alam)
{
    if (!out || !ncell || !dx) return fail(CNS_EINVAL, "NULL argument");
    *out = nullptr;
    if (max_grid_size < NG) return fail(CNS_EINVAL, "max_grid_size %d is below %d", max_grid_size, NG);
    int nb[3];
    int64_t nbox = 1;
    for (int d = 0; d < 3; d++) {
        if (ncell[d] < NG || ncell[d] % max_grid_size)
            return fail(CNS_EINVAL, "extent %d along %d is not a positive multiple of max_grid_size %d",
                        ncell[d], d, max_grid_size);
        if (!(dx[d] > 0.0)) return fail(CNS_EINVAL, "dx[%d] must be positive", d);
        nb[d] = ncell[d] / max_grid_size;
        nbox *= nb[d];
    }
    if (nbox > 65535) return fail(CNS_EINVAL, "%lld boxes; at most 65535 are supported", (long long)nbox);
    cns_domain *D = new (std::nothrow) cns_domain();
    if (!D) return fail(CNS_ENOMEM, "host allocation failed");
    D->g = make_geom(max_grid_size, max_grid_size, max_grid_size);
    D->bs = box_set(D->g, (int)nbox);
    D->m = max_grid_size;
    for (int d = 0; d < 3; d++) {
        D->nb[d] = nb[d];
        D->dx[d] = dx[d];
        D->dxi[d] = 1.0E0 / dx[d];
    }
    D->cfl = cfl;
    D->eta = eta;
    D->alam = alam;
    void *blk = nullptr;
    if (hipMalloc(&blk, 4 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        delete D;
        return fail(CNS_ENOMEM, "hipMalloc of the domain block failed");
    }
    D->d_acc = (unsigned long long *)blk;
    D->d_dt = (double *)blk + 1;
    D->d_courno = (double *)blk + 2;
    const size_t stage = (size_t)nbox * D->bs.su, wz = (size_t)nbox * D->bs.sw;
    if (hipMalloc((void **)&D->d_stage, stage * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&D->d_wz, wz * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        if (D->d_stage) (void)hipFree(D->d_stage);
        (void)hipFree(blk);
        delete D;
        return fail(CNS_ENOMEM, "hipMalloc of the %zu-byte stage and wz buffers failed",
                    (stage + wz) * sizeof(double));
    }
    *out = D;
    return CNS_OK;
}

int cns_domain_destroy(cns_domain *D)
{
    if (!D) return CNS_OK;
    (void)hipFree(D->d_acc);
    (void)hipFree(D->d_stage);
    (void)hipFree(D->d_wz);
    delete D;
    return CNS_OK;
}

int cns_domain_boxes(const cns_domain *D, int nb[3], int *nbox)
{
    if (!D || !nb || !nbox) return fail(CNS_EINVAL, "NULL argument");
    for (int d = 0; d < 3; d++) nb[d] = D->nb[d];
    *nbox = D->bs.n;
    return CNS_OK;
}

int cns_domain_box(const cns_domain *D, int n, int lo[3], int hi[3])
{
    if (!D || !lo || !hi) return fail(CNS_EINVAL, "NULL argument");
    if (n < 0 || n >= D->bs.n) return fail(CNS_EINVAL, "box %d of %d", n, D->bs.n);
    const int I[3] = {n % D->nb[0], (n / D->nb[0]) % D->nb[1], n / (D->nb[0] * D->nb[1])};
    for (int d = 0; d < 3; d++) {
        lo[d] = I[d] * D->m;
        hi[d] = lo[d] + D->m - 1;
    }
    return CNS_OK;
}

size_t cns_domain_array_len(const cns_domain *D, int ncomp, int ghosted)
{
    if (!D || ncomp < 1) return 0;
    return (size_t)D->bs.n * ncomp * (size_t)(ghosted ? D->g.cg : D->g.c);
}

int cns_domain_fill_boundary(cns_domain *D, double *U)
{
    if (!D || !U) return fail(CNS_EINVAL, "NULL argument");
    CNS_HIPCHK(launch_domain_fill(D->g, D->bs, D->nb, U, nullptr, 0u, current_stream()));
    return CNS_OK;
}

int cns_domain_advance(cns_domain *D, double *U, double *Unew, double *Q, double *Dd, double *F,
                       double *dt)
{
    if (!D || !U || !Unew || !Q || !Dd || !F || !dt) return fail(CNS_EINVAL, "NULL argument");
    for (const void *p : {(const void *)U, (const void *)Unew, (const void *)Q, (const void *)Dd, (const void *)F})
        if (!pipe_aligned(p)) return fail(CNS_EINVAL, "domain arrays must be 16-byte aligned");
    const Geom &g = D->g;
    const BoxSet &bs = D->bs;
    const double *dxi = D->dxi;
    double *T = D->d_stage, *WZ = D->d_wz;
    hipStream_t s = current_stream();
    const unsigned uvwt = (1u << QU) | (1u << QV) | (1u << QW) | (1u << QTEMP);
    // stage 1 (fortran90/advance.f90:81-262): box fill, ctoprim + Courant over every box
    CNS_HIPCHK(launch_domain_fill(g, bs, D->nb, U, nullptr, 0u, s));
    CNS_HIPCHK(launch_courno_reset(D->d_acc, 1.0E-50, s));
    CNS_HIPCHK(launch_ctoprim_wz(g, U, Q, WZ, D->d_acc, D->dx[0], D->dx[1], D->dx[2], dxi[2], false, s, bs));
    CNS_HIPCHK(launch_courno_to_dt(D->d_acc, D->cfl, D->d_dt, D->d_courno, s));
    CNS_HIPCHK(launch_diffterm_pipe(g, Q, WZ, Dd, dxi[0], dxi[1], dxi[2], D->eta, D->alam, s, false, bs));
    CNS_HIPCHK(launch_hypterm_cons(g, U, F, dxi[0], dxi[1], dxi[2], Dd, U, T, D->d_dt, 1, false, Q, 1, s, bs));
    // stage 2 (:264-336): the stage kernel stored u, v, w, T of U^{1/3}; the fill
    // takes them across boxes with the state
    CNS_HIPCHK(launch_domain_fill(g, bs, D->nb, T, Q, uvwt, s));
    CNS_HIPCHK(launch_wz(g, Q, WZ, dxi[2], s, bs));
    CNS_HIPCHK(launch_diffterm_pipe(g, Q, WZ, Dd, dxi[0], dxi[1], dxi[2], D->eta, D->alam, s, false, bs));
    CNS_HIPCHK(launch_hypterm_cons(g, T, F, dxi[0], dxi[1], dxi[2], Dd, U, Unew, D->d_dt, 2, false, Q, 2, s, bs));
    // stage 3 (:338-410)
    CNS_HIPCHK(launch_domain_fill(g, bs, D->nb, Unew, Q, (1u << NQ) - 1u, s));
    CNS_HIPCHK(launch_wz(g, Q, WZ, dxi[2], s, bs));
    CNS_HIPCHK(launch_diffterm_pipe(g, Q, WZ, Dd, dxi[0], dxi[1], dxi[2], D->eta, D->alam, s, true, bs));
    CNS_HIPCHK(launch_hypterm_cons(g, Unew, F, dxi[0], dxi[1], dxi[2], Dd, U, U, D->d_dt, 3, true, nullptr, 0, s, bs));
    double h[2];
    CNS_HIPCHK(hipMemcpyAsync(h, D->d_dt, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CNS_HIPCHK(hipStreamSynchronize(s));
    *dt = h[0];
    if (!std::isfinite(h[1]) || !std::isfinite(h[0]))
        return fail(CNS_ENUMERIC, "non-finite Courant number %g (dt %g)", h[1], h[0]);
    return CNS_OK;
}

}  // extern "C"
