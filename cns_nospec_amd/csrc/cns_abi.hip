// cns_abi.hip -- the extern "C" entry points of libcns_hip.so (declared in include/cns.h).
//
// Argument checking, stream selection, error reporting and the composition of
// the RK3 step.  All device work is enqueued on the calling thread's current
// stream; the library never calls exit() and never allocates on the step path.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "cns.h"
#include "cns_host.h"
#include "cns_kernels.h"

namespace cns {

namespace {
thread_local hipStream_t t_stream = nullptr;
thread_local std::string t_err;
}  // namespace

hipStream_t current_stream() { return t_stream; }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

}  // namespace cns

using namespace cns;

#define HIPCHK(call) CNS_HIPCHK(call)

namespace {

int g_stencil_impl = CNS_IMPL_TILED;

// Geometry from the ghost-shifted bounds of the reference (lo = ng, hi = n-1+ng).
int geom_from_bounds(const int lo[3], const int hi[3], int ng, Geom *g)
{
    if (!lo || !hi) return fail(CNS_EINVAL, "lo/hi must not be NULL");
    if (ng != NG) return fail(CNS_EINVAL, "ng = %d, only ng = %d is supported", ng, NG);
    int n[3];
    for (int d = 0; d < 3; d++) {
        if (lo[d] != ng) return fail(CNS_EINVAL, "lo[%d] = %d, expected ng = %d", d, lo[d], ng);
        n[d] = hi[d] - lo[d] + 1;
        if (n[d] < NG) return fail(CNS_EINVAL, "extent %d along %d is below %d", n[d], d, NG);
    }
    *g = make_geom(n[0], n[1], n[2]);
    return CNS_OK;
}

// Caller-owned device workspace of the calling thread (cns_set_scratch): when it is
// large enough for a call, cns_ctoprim's Courant slot and cns_diffterm's wz field live
// there and the library allocates nothing; otherwise the per-thread buffers below are
// allocated on first use.
thread_local void *t_ws = nullptr;
thread_local size_t t_ws_bytes = 0;
thread_local int t_ws_device = -1;

// The workspace's first `bytes` bytes, when the caller set one on this device that holds them.
void *caller_workspace(size_t bytes)
{
    int dev = -1;
    if (!t_ws || t_ws_bytes < bytes || hipGetDevice(&dev) != hipSuccess || dev != t_ws_device) return nullptr;
    return t_ws;
}

// One Courant accumulator per (thread, device) for cns_ctoprim's host-courno form.
struct TlsSlots {
    unsigned long long *acc[64] = {};
    ~TlsSlots()
    {
        for (auto *p : acc)
            if (p) (void)hipFree(p);
    }
};
thread_local TlsSlots t_slots;

int thread_acc(unsigned long long **out)
{
    if (void *w = caller_workspace(sizeof(unsigned long long))) {
        *out = (unsigned long long *)w;
        return CNS_OK;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(CNS_EINVAL, "device %d out of range", dev);
    if (!t_slots.acc[dev]) {
        if (hipMalloc((void **)&t_slots.acc[dev], sizeof(unsigned long long)) != hipSuccess)
            return fail(CNS_ENOMEM, "hipMalloc of the Courant slot failed");
    }
    *out = t_slots.acc[dev];
    return CNS_OK;
}

// Device scratch of the free-standing cns_diffterm (the wz field of the pipelined
// kernel): one grow-only buffer per (thread, device), reused by later calls.
struct TlsScratch {
    double *p[64] = {};
    size_t n[64] = {};
    ~TlsScratch()
    {
        for (auto *q : p)
            if (q) (void)hipFree(q);
    }
};
thread_local TlsScratch t_scratch;

int thread_scratch(size_t count, double **out)
{
    if (void *w = caller_workspace(count * sizeof(double))) {
        *out = (double *)w;
        return CNS_OK;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(CNS_EINVAL, "device %d out of range", dev);
    if (t_scratch.n[dev] < count) {
        if (t_scratch.p[dev]) {
            HIPCHK(hipStreamSynchronize(t_stream));
            HIPCHK(hipFree(t_scratch.p[dev]));
            t_scratch.p[dev] = nullptr;
            t_scratch.n[dev] = 0;
        }
        if (hipMalloc((void **)&t_scratch.p[dev], count * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            t_scratch.p[dev] = nullptr;
            return fail(CNS_ENOMEM, "hipMalloc of %zu bytes of diffterm scratch failed", count * sizeof(double));
        }
        t_scratch.n[dev] = count;
    }
    *out = t_scratch.p[dev];
    return CNS_OK;
}

// Scratch of cns_diffterm_noalloc: the wz field of the pipelined diffterm (one
// ghosted component) for boxes whose ghosted component holds at most
// CNS_NOALLOC_SCRATCH_DOUBLES values.  Each GPU the library runs on holds its own
// copy until the process ends; every call on that GPU shares it.  In host code the
// name is only a handle: kernels get the per-device address from
// hipGetSymbolAddress (noalloc_scratch below).
__device__ __align__(256) double scratch[CNS_NOALLOC_SCRATCH_DOUBLES];

std::mutex g_scratch_mu;
double *g_scratch_dev[64] = {};

// Device address of `scratch` on the current device, looked up once per device.
int noalloc_scratch(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(CNS_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(CNS_EHIP, "hipGetSymbolAddress of the no-alloc scratch failed: %s",
                        hipGetErrorString(e));
        }
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return CNS_OK;
}

// The tiled implementation runs the DMA-pipelined kernels (cns_stencil_pipe.hip)
// when every array is 16-byte aligned, else the register-staged ones
// (cns_stencil.hip).  CNS_PIPE=0 in the environment selects the latter always.
bool pipe_on(std::initializer_list<const void *> ptrs)
{
    const char *e = getenv("CNS_PIPE");
    if (e && atoi(e) == 0) return false;
    for (const void *p : ptrs)
        if (p && !pipe_aligned(p)) return false;
    return true;
}

// The pipelined fused step stores each stage's primitives from the stage kernel
// (no ctoprim pass in stages 2 and 3); CNS_STAGE_Q=0 selects a ctoprim pass per stage.
bool stage_q_on()
{
    const char *e = getenv("CNS_STAGE_Q");
    return !(e && atoi(e) == 0);
}

hipError_t stencil_hypterm(const Geom &g, const double *cons, const double *q, double *flux,
                           const double dxi[3], hipStream_t s)
{
    if (g_stencil_impl == CNS_IMPL_DIRECT)
        return launch_hypterm_direct(g, cons, q, flux, dxi[0], dxi[1], dxi[2], s);
    if (pipe_on({cons, q, flux}))
        return launch_hypterm_pipe(g, cons, q, flux, dxi[0], dxi[1], dxi[2], nullptr, nullptr,
                                   nullptr, nullptr, 0, true, s);
    return launch_hypterm_tiled(g, cons, q, flux, dxi[0], dxi[1], dxi[2], s);
}

// write_f: whether F must hold this stage's flux afterwards (only the last stage's
// F is an output of the step; the register-staged kernel always stores it).
hipError_t stencil_hypterm_rk(const Geom &g, const double *cons, const double *q, double *flux,
                              const double dxi[3], const double *d, const double *base, double *out,
                              const double *dt, int phase, bool write_f, hipStream_t s)
{
    if (pipe_on({cons, q, flux, d, base, out}))
        return launch_hypterm_pipe(g, cons, q, flux, dxi[0], dxi[1], dxi[2], d, base, out, dt, phase,
                                   write_f, s);
    return launch_hypterm_rk(g, cons, q, flux, dxi[0], dxi[1], dxi[2], d, base, out, dt, phase, s);
}

// wz: scratch of one ghosted component (g.cg doubles), or null when the caller has
// none (the register-staged kernel needs none).
hipError_t stencil_diffterm(const Geom &g, const double *q, double *difflux, const double dxi[3],
                            double eta, double alam, double *wz, hipStream_t s)
{
    if (g_stencil_impl == CNS_IMPL_DIRECT)
        return launch_diffterm_direct(g, q, difflux, dxi[0], dxi[1], dxi[2], eta, alam, s);
    if (wz && pipe_on({q, difflux, wz})) {
        hipError_t e = launch_wz(g, q, wz, dxi[2], s);
        if (e != hipSuccess) return e;
        return launch_diffterm_pipe(g, q, wz, difflux, dxi[0], dxi[1], dxi[2], eta, alam, s);
    }
    return launch_diffterm_tiled(g, q, difflux, dxi[0], dxi[1], dxi[2], eta, alam, s);
}

}  // namespace

// Live per-kernel timing of the step: a pair of HIP events around each launch,
// recorded on the launch stream and folded into totals on read.
struct Profiler {
    bool on = false;
    struct Pair { hipEvent_t a, b; int kind; };
    std::vector<Pair> pool;
    size_t used = 0;
    double ms[CNS_PROF_NKINDS] = {};
    long long n[CNS_PROF_NKINDS] = {};

    hipError_t begin(int kind, hipStream_t s, Pair **out)
    {
        *out = nullptr;
        if (!on) return hipSuccess;
        if (used == pool.size()) {
            Pair p{};
            hipError_t e = hipEventCreate(&p.a);
            if (e != hipSuccess) return e;
            if ((e = hipEventCreate(&p.b)) != hipSuccess) return e;
            pool.push_back(p);
        }
        Pair *p = &pool[used++];
        p->kind = kind;
        *out = p;
        return hipEventRecord(p->a, s);
    }
    hipError_t end(Pair *p, hipStream_t s) { return p ? hipEventRecord(p->b, s) : hipSuccess; }
    hipError_t fold()
    {
        for (size_t i = 0; i < used; i++) {
            float t = 0.f;
            hipError_t e = hipEventSynchronize(pool[i].b);
            if (e != hipSuccess) return e;
            if ((e = hipEventElapsedTime(&t, pool[i].a, pool[i].b)) != hipSuccess) return e;
            ms[pool[i].kind] += t;
            n[pool[i].kind] += 1;
        }
        used = 0;
        return hipSuccess;
    }
    void reset()
    {
        used = 0;
        for (int k = 0; k < CNS_PROF_NKINDS; k++) { ms[k] = 0.0; n[k] = 0; }
    }
    ~Profiler()
    {
        for (auto &p : pool) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
};

struct cns_ctx {
    Geom g;
    double dx[3], dxi[3];
    double cfl, eta, alam;
    int device;
    // one device block: [0] Courant accumulator bits, [1] dt, [2] courno
    unsigned long long *d_acc;
    double *d_dt, *d_courno;
    // Stage buffer of the fused step (ghosted, 5 comps): holds U^{1/3}, so that the
    // fused hypterm+update never writes an array it reads with halos.  Null when the
    // unfused composition is selected.
    double *d_stage;
    // wz field of the pipelined diffterm (one ghosted component).
    double *d_wz;
    Profiler prof;
};

// Runs one launch, timed when the context's profiler is on.
#define TIMED(ctx, kind, s, call)                                 \
    do {                                                          \
        Profiler::Pair *pp_ = nullptr;                            \
        HIPCHK((ctx)->prof.begin((kind), (s), &pp_));             \
        HIPCHK(call);                                             \
        HIPCHK((ctx)->prof.end(pp_, (s)));                        \
        if ((ctx)->prof.used >= 8192) HIPCHK((ctx)->prof.fold()); \
    } while (0)

extern "C" {

const char *cns_last_error_string(void) { return t_err.c_str(); }

int cns_set_stream(void *hip_stream)
{
    t_stream = (hipStream_t)hip_stream;
    return CNS_OK;
}

void *cns_get_stream(void) { return (void *)t_stream; }

int cns_set_scratch(void *dev_ptr, size_t bytes)
{
    if (!dev_ptr && bytes) return fail(CNS_EINVAL, "NULL workspace with a nonzero size");
    if (dev_ptr && ((uintptr_t)dev_ptr % 256) != 0)
        return fail(CNS_EINVAL, "workspace must be 256-byte aligned");
    int dev = -1;
    if (dev_ptr) HIPCHK(hipGetDevice(&dev));
    t_ws = dev_ptr;
    t_ws_bytes = dev_ptr ? bytes : 0;
    t_ws_device = dev;
    return CNS_OK;
}

size_t cns_scratch_bytes(const int lo[3], const int hi[3], int ng)
{
    Geom g;
    if (geom_from_bounds(lo, hi, ng, &g)) return 0;
    return (size_t)g.cg * sizeof(double);  // the wz field; the Courant slot fits in it
}

int cns_set_stencil_impl(int impl)
{
    if (impl != CNS_IMPL_TILED && impl != CNS_IMPL_DIRECT)
        return fail(CNS_EINVAL, "unknown stencil implementation %d", impl);
    g_stencil_impl = impl;
    return CNS_OK;
}

int cns_get_stencil_impl(void) { return g_stencil_impl; }

int cns_fill_boundary(double *u, const int dim[3], const int dim_g[3])
{
    if (!u || !dim || !dim_g) return fail(CNS_EINVAL, "NULL argument");
    for (int d = 0; d < 3; d++) {
        if (dim[d] < NG) return fail(CNS_EINVAL, "extent %d along %d is below %d", dim[d], d, NG);
        if (dim_g[d] != dim[d] + 2 * NG)
            return fail(CNS_EINVAL, "dim_g[%d] = %d, expected %d", d, dim_g[d], dim[d] + 2 * NG);
    }
    Geom g = make_geom(dim[0], dim[1], dim[2]);
    HIPCHK(launch_fill_boundary(g, u, t_stream));
    return CNS_OK;
}

int cns_ctoprim(const int lo[3], const int hi[3], const double *u, double *q, const double dx[3],
                int ng, double *courno)
{
    Geom g;
    int rc = geom_from_bounds(lo, hi, ng, &g);
    if (rc) return rc;
    if (!u || !q || !dx) return fail(CNS_EINVAL, "NULL argument");
    if (!courno) {
        HIPCHK(launch_ctoprim(g, u, q, nullptr, dx[0], dx[1], dx[2], t_stream));
        return CNS_OK;
    }
    unsigned long long *acc = nullptr;
    if ((rc = thread_acc(&acc))) return rc;
    HIPCHK(launch_courno_reset(acc, *courno, t_stream));
    HIPCHK(launch_ctoprim(g, u, q, acc, dx[0], dx[1], dx[2], t_stream));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, acc, sizeof bits, hipMemcpyDeviceToHost, t_stream));
    HIPCHK(hipStreamSynchronize(t_stream));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    // courno = MAX(MAX(courmx,courmy), MAX(courmz,courno)) (cpu/ctoprim.cpp:69)
    *courno = m > *courno ? m : *courno;
    if (!std::isfinite(m)) *courno = m;
    return CNS_OK;
}

int cns_hypterm(const int lo[3], const int hi[3], int ng, const double dx[3], const double *cons,
                const double *q, double *flux)
{
    Geom g;
    int rc = geom_from_bounds(lo, hi, ng, &g);
    if (rc) return rc;
    if (!dx || !cons || !q || !flux) return fail(CNS_EINVAL, "NULL argument");
    // dxinv(i) = 1/dx(i) (cpu/hypterm.cpp:25-26)
    double dxi[3] = {1.0E0 / dx[0], 1.0E0 / dx[1], 1.0E0 / dx[2]};
    HIPCHK(stencil_hypterm(g, cons, q, flux, dxi, t_stream));
    return CNS_OK;
}

int cns_diffterm(const int lo[3], const int hi[3], int ng, const double dx[3], const double *q,
                 double *difflux, double eta, double alam)
{
    Geom g;
    int rc = geom_from_bounds(lo, hi, ng, &g);
    if (rc) return rc;
    if (!dx || !q || !difflux) return fail(CNS_EINVAL, "NULL argument");
    double dxi[3] = {1.0E0 / dx[0], 1.0E0 / dx[1], 1.0E0 / dx[2]};  // cpu/diffterm.cpp:54-57
    double *wz = nullptr;
    if (g_stencil_impl == CNS_IMPL_TILED && pipe_on({q, difflux})) {
        rc = thread_scratch((size_t)g.cg, &wz);
        if (rc) return rc;
    }
    HIPCHK(stencil_diffterm(g, q, difflux, dxi, eta, alam, wz, t_stream));
    return CNS_OK;
}

int cns_diffterm_noalloc(const int lo[3], const int hi[3], int ng, const double dx[3],
                         const double *q, double *difflux, double eta, double alam)
{
    Geom g;
    int rc = geom_from_bounds(lo, hi, ng, &g);
    if (rc) return rc;
    if (!dx || !q || !difflux) return fail(CNS_EINVAL, "NULL argument");
    if ((size_t)g.cg > (size_t)CNS_NOALLOC_SCRATCH_DOUBLES)
        return fail(CNS_EINVAL,
                    "box %dx%dx%d: its ghosted component holds %lld values, above the no-alloc "
                    "scratch of %d (use cns_diffterm)",
                    g.n0, g.n1, g.n2, (long long)g.cg, CNS_NOALLOC_SCRATCH_DOUBLES);
    double dxi[3] = {1.0E0 / dx[0], 1.0E0 / dx[1], 1.0E0 / dx[2]};  // cpu/diffterm.cpp:54-57
    double *wz = nullptr;
    if (g_stencil_impl == CNS_IMPL_TILED && pipe_on({q, difflux})) {
        rc = noalloc_scratch(&wz);
        if (rc) return rc;
    }
    HIPCHK(stencil_diffterm(g, q, difflux, dxi, eta, alam, wz, t_stream));
    return CNS_OK;
}

int cns_rk_update(int phase, const int lo[3], const int hi[3], int ng, double *u, double *unew,
                  const double *d, const double *f, double dt)
{
    Geom g;
    int rc = geom_from_bounds(lo, hi, ng, &g);
    if (rc) return rc;
    if (phase < 1 || phase > 3) return fail(CNS_EINVAL, "phase %d not in 1..3", phase);
    if (!u || !unew || !d || !f) return fail(CNS_EINVAL, "NULL argument");
    HIPCHK(launch_rk_update(g, phase, u, unew, d, f, nullptr, dt, t_stream));
    return CNS_OK;
}

int cns_ctx_create(cns_ctx **ctx, const int ncells[3], const double dx[3], double cfl, double eta,
                   double alam)
{
    if (!ctx || !ncells || !dx) return fail(CNS_EINVAL, "NULL argument");
    *ctx = nullptr;
    for (int d = 0; d < 3; d++) {
        if (ncells[d] < NG) return fail(CNS_EINVAL, "extent %d along %d is below %d", ncells[d], d, NG);
        if (!(dx[d] > 0.0)) return fail(CNS_EINVAL, "dx[%d] must be positive", d);
    }
    cns_ctx *c = new (std::nothrow) cns_ctx();
    if (!c) return fail(CNS_ENOMEM, "host allocation failed");
    c->g = make_geom(ncells[0], ncells[1], ncells[2]);
    for (int d = 0; d < 3; d++) {
        c->dx[d] = dx[d];
        c->dxi[d] = 1.0E0 / dx[d];
    }
    c->cfl = cfl;
    c->eta = eta;
    c->alam = alam;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(CNS_EHIP, "hipGetDevice failed");
    }
    void *blk = nullptr;
    if (hipMalloc(&blk, 4 * sizeof(double)) != hipSuccess) {
        delete c;
        return fail(CNS_ENOMEM, "hipMalloc of the context block failed");
    }
    c->d_acc = (unsigned long long *)blk;
    c->d_dt = (double *)blk + 1;
    c->d_courno = (double *)blk + 2;
    c->d_stage = nullptr;
    c->d_wz = nullptr;
    if (hipMalloc((void **)&c->d_stage, (size_t)NC * c->g.cg * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&c->d_wz, (size_t)c->g.cg * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        if (c->d_stage) (void)hipFree(c->d_stage);
        (void)hipFree(blk);
        delete c;
        return fail(CNS_ENOMEM, "hipMalloc of the %zu-byte stage and wz buffers failed",
                    (size_t)(NC + 1) * c->g.cg * sizeof(double));
    }
    *ctx = c;
    return CNS_OK;
}

int cns_ctx_set_fused(cns_ctx *ctx, int fused)
{
    if (!ctx) return fail(CNS_EINVAL, "NULL context");
    if (!fused && ctx->d_stage) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(ctx->d_stage));
        ctx->d_stage = nullptr;
    } else if (fused && !ctx->d_stage) {
        if (hipMalloc((void **)&ctx->d_stage, (size_t)NC * ctx->g.cg * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            ctx->d_stage = nullptr;
            return fail(CNS_ENOMEM, "hipMalloc of the stage buffer failed");
        }
    }
    return CNS_OK;
}

int cns_ctx_destroy(cns_ctx *ctx)
{
    if (!ctx) return CNS_OK;
    (void)hipFree(ctx->d_acc);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->d_wz) (void)hipFree(ctx->d_wz);
    delete ctx;
    return CNS_OK;
}

int cns_advance_async(cns_ctx *ctx, double *U, double *Unew, double *Q, double *D, double *F,
                      double *dt_dev, double *courno_dev)
{
    if (!ctx || !U || !Unew || !Q || !D || !F) return fail(CNS_EINVAL, "NULL argument");
    const Geom &g = ctx->g;
    hipStream_t s = t_stream;
    double *dt = dt_dev ? dt_dev : ctx->d_dt;
    if (ctx->d_stage) {
        // Fused composition: each stage is fill, ctoprim, diffterm, and hypterm with the
        // RK update of its cell applied in the same pass (F is still written).  Stage
        // results go to a buffer the stage does not read with halos:
        //   stage 1: U -> stage (U^{1/3});  stage 2: stage -> Unew (U^{2/3});
        //   stage 3: Unew -> U (U^{n+1}).
        // Every ABI array ends as the reference leaves it (cpu/advance.cpp:13-154).
        double *T = ctx->d_stage;
        if (pipe_on({U, Unew, Q, D, F, T, ctx->d_wz})) {
            // q is ctoprim(state) here, so the stage kernel derives u, v, w, p from the
            // conserved fields itself, and the first two stages' ctoprim stores only the
            // u, v, w, T that diffterm reads; the third stores all of Q, as the
            // reference leaves it.
            // ctoprim also writes the wz field diffterm reads (one z-marching pass).
            const double *dxi = ctx->dxi;
            double *WZ = ctx->d_wz;
            if (stage_q_on()) {
                // Stages 1 and 2 also store ctoprim of their result (interior cells) into
                // Q, so stages 2 and 3 skip the ctoprim pass: the periodic fill copies
                // those cells' Q into the ghosts with the state's, and a wz pass reads w.
                // diffterm's rho row (0.0) is stored by stage 3 only; nothing reads it.
                const unsigned uvwt = (1u << QU) | (1u << QV) | (1u << QW) | (1u << QTEMP);
                TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, U, s));
                HIPCHK(launch_courno_reset(ctx->d_acc, 1.0E-50, s));
                TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim_wz(g, U, Q, WZ, ctx->d_acc, ctx->dx[0], ctx->dx[1], ctx->dx[2], dxi[2], false, s));
                HIPCHK(launch_courno_to_dt(ctx->d_acc, ctx->cfl, dt, courno_dev ? courno_dev : ctx->d_courno, s));
                TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s, false));
                TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, U, F, dxi[0], dxi[1], dxi[2], D, U, T, dt, 1, false, Q, 1, s));
                TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary_q(g, T, Q, uvwt, s));
                TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_wz(g, Q, WZ, dxi[2], s));
                TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s, false));
                TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, T, F, dxi[0], dxi[1], dxi[2], D, U, Unew, dt, 2, false, Q, 2, s));
                TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary_q(g, Unew, Q, (1u << NQ) - 1u, s));
                TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_wz(g, Q, WZ, dxi[2], s));
                TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s, true));
                TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, Unew, F, dxi[0], dxi[1], dxi[2], D, U, U, dt, 3, true, nullptr, 0, s));
                return CNS_OK;
            }
            TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, U, s));
            HIPCHK(launch_courno_reset(ctx->d_acc, 1.0E-50, s));
            TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim_wz(g, U, Q, WZ, ctx->d_acc, ctx->dx[0], ctx->dx[1], ctx->dx[2], dxi[2], false, s));
            HIPCHK(launch_courno_to_dt(ctx->d_acc, ctx->cfl, dt, courno_dev ? courno_dev : ctx->d_courno, s));
            TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s));
            TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, U, F, ctx->dxi[0], ctx->dxi[1], ctx->dxi[2], D, U, T, dt, 1, false, nullptr, 0, s));
            TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, T, s));
            TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim_wz(g, T, Q, WZ, nullptr, 0.0, 0.0, 0.0, dxi[2], false, s));
            TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s));
            TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, T, F, ctx->dxi[0], ctx->dxi[1], ctx->dxi[2], D, U, Unew, dt, 2, false, nullptr, 0, s));
            TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, Unew, s));
            TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim_wz(g, Unew, Q, WZ, nullptr, 0.0, 0.0, 0.0, dxi[2], true, s));
            TIMED(ctx, CNS_PROF_DIFFTERM, s, launch_diffterm_pipe(g, Q, WZ, D, dxi[0], dxi[1], dxi[2], ctx->eta, ctx->alam, s));
            TIMED(ctx, CNS_PROF_HYPTERM_RK, s, launch_hypterm_cons(g, Unew, F, ctx->dxi[0], ctx->dxi[1], ctx->dxi[2], D, U, U, dt, 3, true, nullptr, 0, s));
            return CNS_OK;
        }
        TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, U, s));
        HIPCHK(launch_courno_reset(ctx->d_acc, 1.0E-50, s));
        TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, U, Q, ctx->d_acc, ctx->dx[0], ctx->dx[1], ctx->dx[2], s));
        HIPCHK(launch_courno_to_dt(ctx->d_acc, ctx->cfl, dt, courno_dev ? courno_dev : ctx->d_courno, s));
        TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
        TIMED(ctx, CNS_PROF_HYPTERM_RK, s, stencil_hypterm_rk(g, U, Q, F, ctx->dxi, D, U, T, dt, 1, false, s));

        TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, T, s));
        TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, T, Q, nullptr, 0.0, 0.0, 0.0, s));
        TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
        TIMED(ctx, CNS_PROF_HYPTERM_RK, s, stencil_hypterm_rk(g, T, Q, F, ctx->dxi, D, U, Unew, dt, 2, false, s));

        TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, Unew, s));
        TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, Unew, Q, nullptr, 0.0, 0.0, 0.0, s));
        TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
        TIMED(ctx, CNS_PROF_HYPTERM_RK, s, stencil_hypterm_rk(g, Unew, Q, F, ctx->dxi, D, U, U, dt, 3, true, s));
        return CNS_OK;
    }
    // Unfused composition, kernel for kernel as the reference.
    // cpu/advance.cpp:45-85: stage 1
    TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, U, s));
    HIPCHK(launch_courno_reset(ctx->d_acc, 1.0E-50, s));
    TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, U, Q, ctx->d_acc, ctx->dx[0], ctx->dx[1], ctx->dx[2], s));
    HIPCHK(launch_courno_to_dt(ctx->d_acc, ctx->cfl, dt, courno_dev ? courno_dev : ctx->d_courno, s));
    TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
    TIMED(ctx, CNS_PROF_HYPTERM, s, stencil_hypterm(g, U, Q, F, ctx->dxi, s));
    TIMED(ctx, CNS_PROF_RK, s, launch_rk_update(g, 1, U, Unew, D, F, dt, 0.0, s));
    // :90-119: stage 2
    TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, Unew, s));
    TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, Unew, Q, nullptr, 0.0, 0.0, 0.0, s));
    TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
    TIMED(ctx, CNS_PROF_HYPTERM, s, stencil_hypterm(g, Unew, Q, F, ctx->dxi, s));
    TIMED(ctx, CNS_PROF_RK, s, launch_rk_update(g, 2, U, Unew, D, F, dt, 0.0, s));
    // :124-153: stage 3
    TIMED(ctx, CNS_PROF_FILL, s, launch_fill_boundary(g, Unew, s));
    TIMED(ctx, CNS_PROF_CTOPRIM, s, launch_ctoprim(g, Unew, Q, nullptr, 0.0, 0.0, 0.0, s));
    TIMED(ctx, CNS_PROF_DIFFTERM, s, stencil_diffterm(g, Q, D, ctx->dxi, ctx->eta, ctx->alam, ctx->d_wz, s));
    TIMED(ctx, CNS_PROF_HYPTERM, s, stencil_hypterm(g, Unew, Q, F, ctx->dxi, s));
    TIMED(ctx, CNS_PROF_RK, s, launch_rk_update(g, 3, U, Unew, D, F, dt, 0.0, s));
    return CNS_OK;
}

int cns_advance(cns_ctx *ctx, double *U, double *Unew, double *Q, double *D, double *F, double *dt)
{
    if (!dt) return fail(CNS_EINVAL, "NULL dt");
    int rc = cns_advance_async(ctx, U, Unew, Q, D, F, nullptr, nullptr);
    if (rc) return rc;
    double h[2];
    HIPCHK(hipMemcpyAsync(h, ctx->d_dt, 2 * sizeof(double), hipMemcpyDeviceToHost, t_stream));
    HIPCHK(hipStreamSynchronize(t_stream));
    *dt = h[0];
    if (!std::isfinite(h[1]) || !std::isfinite(h[0]))
        return fail(CNS_ENUMERIC, "non-finite Courant number %g (dt %g)", h[1], h[0]);
    return CNS_OK;
}

int cns_ctx_profile(cns_ctx *ctx, int enable)
{
    if (!ctx) return fail(CNS_EINVAL, "NULL context");
    HIPCHK(ctx->prof.fold());
    ctx->prof.reset();
    ctx->prof.on = enable != 0;
    return CNS_OK;
}

int cns_ctx_profile_read(cns_ctx *ctx, double ms_total[CNS_PROF_NKINDS],
                         long long launches[CNS_PROF_NKINDS])
{
    if (!ctx || !ms_total || !launches) return fail(CNS_EINVAL, "NULL argument");
    HIPCHK(ctx->prof.fold());
    for (int k = 0; k < CNS_PROF_NKINDS; k++) {
        ms_total[k] = ctx->prof.ms[k];
        launches[k] = ctx->prof.n[k];
    }
    return CNS_OK;
}

}  // extern "C"
