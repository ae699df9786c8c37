// cns_batch_engine.hip -- host side of the batched engine (include/cns_batch.h).
//
// One device pool of the budget size is reserved at creation and carved by a
// first-fit allocator, so admission is a pure bookkeeping decision (a box that
// does not fit is refused, never an out-of-memory fault) and nothing is
// allocated while stepping.  Box records (BoxDev) live in a device array indexed
// by box id.  The work list holds every resident box's tiles, grouped by tile-
// width class; a step (3 RK stages x {wz, diffterm, hypterm+update} per class, then
// the dt finalize) is captured once into a hipGraph and replayed until the
// resident set or the tile table changes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "cns_batch.h"
#include "cns_batch_dev.h"
#include "cns_host.h"

using namespace cns;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with coalescing on release.
struct Pool {
    char *base = nullptr;
    size_t size = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> length

    void init(char *b, size_t n)
    {
        base = b;
        size = n;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(size_t n, size_t *off)
    {
        n = align_up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, roff = it->first + n;
                free_.erase(it);
                if (rest) free_[roff] = rest;
                used += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align_up(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct BoxHost {
    bool live = false;
    int e = 0;
    size_t off = 0, bytes = 0;
};

int width_class(int w) { return w == 16 ? 0 : (w == 32 ? 1 : (w == 64 ? 2 : -1)); }

// Default tile: 32 x 8 columns for every edge, so all boxes share one width class and
// an RK stage is two launches; z runs of 16 planes.  On the mixed request stream (edges
// 24-128) this steps 8 % faster in FP32 than per-edge widths chosen by lane efficiency
// or by timing each edge alone (4.58 vs 4.24 G cell-steps/s; FP64 2.46 vs 2.44 G), and
// runs of 16 planes beat runs of 8 or 32 by 2-3 %.
std::pair<int, int> default_tile(int e) { return {32, std::min(e, 16)}; }

}  // namespace

struct cns_batch {
    int device = 0;
    bool fp32 = false;
    int max_boxes = 0;
    char *pool_mem = nullptr;
    Pool pool;
    BoxDev *d_boxes = nullptr;
    int *d_iota = nullptr;  // d_iota[i] = i, for one-box finalize launches
    int *d_ids = nullptr;   // resident ids
    TileDev *d_tiles = nullptr;
    size_t tile_cap = 0;
    std::vector<BoxHost> hb;
    std::vector<BoxDev> hrec;  // host staging of box records
    std::vector<TileDev> htiles;
    std::vector<int> hids;
    std::map<int, std::pair<int, int>> tune;  // edge -> (width, kz)
    int cls_start[kTileClasses] = {0, 0, 0}, cls_count[kTileClasses] = {0, 0, 0};
    int nresident = 0, ntiles = 0;
    // diffterm with its planes staged by direct-to-LDS loads (and a wz-field pass), the
    // default; CNS_BATCH_DIFF=0 in the environment at creation selects the tile that
    // gathers its halos through registers.  Both give the same bits.
    bool staged_diff = true;
    // FP32 flux, stages 2 and 3: CNS_BATCH_KEEP=1 selects the hypterm tile that keeps
    // its z window's FP64 leading values until they are the centre instead of loading
    // them again (same bits; its 190-200 VGPRs allow 2 workgroups per CU instead of 3,
    // which cost more than the 32 bytes per cell it saves: 4.66 vs 4.73 G cell-steps/s
    // on the request mix).
    bool keep_centre = false;
    // 32-wide tiles: the hypterm + update tile staged by direct-to-LDS loads, FP32 flux
    // and FP64 (CNS_BATCH_HYP=0, CNS_BATCH_HYP64=0 select the tile that gathers its halos
    // through registers; the same bits either way).
    bool staged_hyp = true;
    bool staged_hyp64 = true;
    bool dirty = true;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    std::pair<int, int> tile_of(int e) const
    {
        auto it = tune.find(e);
        return it != tune.end() ? it->second : default_tile(e);
    }
};

namespace {

size_t box_bytes(int e, bool fp32)
{
    const size_t e3 = (size_t)e * e * e;
    const size_t ft = fp32 ? 4 : 8;
    // U, S1, S2 (5 doubles each), Q0..Q2 (5 flux-type values each), D (4), wz (1)
    return 3 * align_up(5 * e3 * 8) + 3 * align_up(5 * e3 * ft) + align_up(4 * e3 * ft) +
           align_up(e3 * ft);
}

void destroy_graph(cns_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Rebuild the work list and re-capture the step graph.
int rebuild(cns_batch *b, hipStream_t s)
{
    CNS_HIPCHK(hipStreamSynchronize(s));  // host staging is about to change
    std::vector<TileDev> cls[kTileClasses];
    b->hids.clear();
    for (int id = 0; id < (int)b->hb.size(); id++) {
        const BoxHost &h = b->hb[id];
        if (!h.live) continue;
        b->hids.push_back(id);
        const auto t = b->tile_of(h.e);
        const int w = t.first, c = width_class(w), kz = std::max(1, std::min(t.second, h.e));
        const int hgt = kBatchThreads / w, ltx = c + 4;
        for (int z0 = 0; z0 < h.e; z0 += kz)
            for (int y0 = 0; y0 < h.e; y0 += hgt)
                for (int x0 = 0; x0 < h.e; x0 += w) {
                    TileDev td;
                    td.box = id;
                    td.x0 = (short)x0;
                    td.y0 = (short)y0;
                    td.z0 = (short)z0;
                    td.nz = (short)std::min(kz, h.e - z0);
                    td.ltx = (short)ltx;
                    td.pad = 0;
                    cls[c].push_back(td);
                }
    }
    b->htiles.clear();
    for (int c = 0; c < kTileClasses; c++) {
        b->cls_start[c] = (int)b->htiles.size();
        b->cls_count[c] = (int)cls[c].size();
        b->htiles.insert(b->htiles.end(), cls[c].begin(), cls[c].end());
    }
    b->nresident = (int)b->hids.size();
    b->ntiles = (int)b->htiles.size();
    if ((size_t)b->ntiles > b->tile_cap) {
        if (b->d_tiles) CNS_HIPCHK(hipFree(b->d_tiles));
        b->d_tiles = nullptr;
        size_t cap = std::max<size_t>((size_t)b->ntiles * 2, 1024);
        if (hipMalloc((void **)&b->d_tiles, cap * sizeof(TileDev)) != hipSuccess) {
            (void)hipGetLastError();
            b->tile_cap = 0;
            return fail(CNS_ENOMEM, "hipMalloc of the tile list (%zu tiles) failed", cap);
        }
        b->tile_cap = cap;
    }
    if (b->ntiles)
        CNS_HIPCHK(hipMemcpyAsync(b->d_tiles, b->htiles.data(), b->ntiles * sizeof(TileDev),
                                  hipMemcpyHostToDevice, s));
    if (b->nresident)
        CNS_HIPCHK(hipMemcpyAsync(b->d_ids, b->hids.data(), b->nresident * sizeof(int),
                                  hipMemcpyHostToDevice, s));
    CNS_HIPCHK(hipStreamSynchronize(s));

    destroy_graph(b);
    if (b->nresident) {
        CNS_HIPCHK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int stage = 1; stage <= 3 && e == hipSuccess; stage++)
            e = launch_batch_stage(b->d_boxes, b->d_tiles, b->cls_start, b->cls_count, stage, b->fp32,
                                   b->staged_diff, b->keep_centre, b->staged_hyp, b->staged_hyp64,
                                   b->cap_stream);
        if (e == hipSuccess) e = launch_batch_finalize(b->d_boxes, b->d_ids, b->nresident, 1, b->cap_stream);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->cap_stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(CNS_EHIP, "capturing the step graph failed: %s",
                        hipGetErrorString(e != hipSuccess ? e : e2));
        }
        b->graph = g;
        CNS_HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->dirty = false;
    return CNS_OK;
}

int check_id(cns_batch *b, int id)
{
    if (!b) return fail(CNS_EINVAL, "NULL engine");
    if (id < 0 || id >= (int)b->hb.size() || !b->hb[id].live)
        return fail(CNS_EINVAL, "box %d is not resident", id);
    return CNS_OK;
}

}  // namespace

extern "C" {

size_t cns_batch_box_bytes(int edge, int flux_fp32) { return box_bytes(edge, flux_fp32 != 0); }

int cns_batch_create(cns_batch **out, size_t hbm_budget, int flux_fp32, int max_boxes)
{
    if (!out) return fail(CNS_EINVAL, "NULL argument");
    *out = nullptr;
    if (max_boxes < 1 || max_boxes > (1 << 20)) return fail(CNS_EINVAL, "max_boxes %d out of range", max_boxes);
    cns_batch *b = new (std::nothrow) cns_batch();
    if (!b) return fail(CNS_ENOMEM, "host allocation failed");
    b->fp32 = flux_fp32 != 0;
    b->max_boxes = max_boxes;
    if (const char *env = getenv("CNS_BATCH_DIFF")) b->staged_diff = atoi(env) != 0;
    if (const char *env = getenv("CNS_BATCH_KEEP")) b->keep_centre = atoi(env) != 0;
    if (const char *env = getenv("CNS_BATCH_HYP")) b->staged_hyp = atoi(env) != 0;
    if (const char *env = getenv("CNS_BATCH_HYP64")) b->staged_hyp64 = atoi(env) != 0;
    int rc = CNS_OK;
    if (hipGetDevice(&b->device) != hipSuccess) rc = fail(CNS_EHIP, "hipGetDevice failed");
    if (!rc && hbm_budget && hipMalloc((void **)&b->pool_mem, hbm_budget) != hipSuccess) {
        (void)hipGetLastError();
        b->pool_mem = nullptr;
        rc = fail(CNS_ENOMEM, "reserving the %zu-byte box pool failed", hbm_budget);
    }
    if (!rc && hipMalloc((void **)&b->d_boxes, max_boxes * sizeof(BoxDev)) != hipSuccess)
        rc = fail(CNS_ENOMEM, "hipMalloc of the box records failed");
    if (!rc && hipMalloc((void **)&b->d_iota, max_boxes * sizeof(int)) != hipSuccess)
        rc = fail(CNS_ENOMEM, "hipMalloc failed");
    if (!rc && hipMalloc((void **)&b->d_ids, max_boxes * sizeof(int)) != hipSuccess)
        rc = fail(CNS_ENOMEM, "hipMalloc failed");
    if (!rc && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(CNS_EHIP, "hipStreamCreate failed");
    if (!rc) {
        std::vector<int> iota(max_boxes);
        for (int i = 0; i < max_boxes; i++) iota[i] = i;
        if (hipMemcpy(b->d_iota, iota.data(), max_boxes * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(CNS_EHIP, "hipMemcpy failed");
    }
    if (rc) {
        cns_batch_destroy(b);
        return rc;
    }
    b->pool.init(b->pool_mem, b->pool_mem ? hbm_budget : 0);
    b->hrec.resize(max_boxes);
    *out = b;
    return CNS_OK;
}

int cns_batch_destroy(cns_batch *b)
{
    if (!b) return CNS_OK;
    (void)hipDeviceSynchronize();
    destroy_graph(b);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_ids) (void)hipFree(b->d_ids);
    if (b->d_iota) (void)hipFree(b->d_iota);
    if (b->d_boxes) (void)hipFree(b->d_boxes);
    if (b->pool_mem) (void)hipFree(b->pool_mem);
    delete b;
    return CNS_OK;
}

int cns_batch_add(cns_batch *b, int edge, double dx, double cfl, double eta, double alam,
                  const double *u0, int *box_id)
{
    if (!b || !u0 || !box_id) return fail(CNS_EINVAL, "NULL argument");
    if (edge < 8 || edge > 512) return fail(CNS_EINVAL, "edge %d outside 8..512", edge);
    if (!(dx > 0.0)) return fail(CNS_EINVAL, "dx must be positive");
    const size_t bytes = box_bytes(edge, b->fp32);
    int id = -1;
    for (int i = 0; i < (int)b->hb.size(); i++)
        if (!b->hb[i].live) {
            id = i;
            break;
        }
    if (id < 0) {
        if ((int)b->hb.size() >= b->max_boxes)
            return fail(CNS_EBUDGET, "engine holds its maximum of %d boxes", b->max_boxes);
        id = (int)b->hb.size();
        b->hb.emplace_back();
    }
    size_t off = 0;
    if (!b->pool.alloc(bytes, &off))
        return fail(CNS_EBUDGET, "box of edge %d (%zu bytes) does not fit: %zu of %zu bytes in use",
                    edge, bytes, b->pool.used, b->pool.size);
    hipStream_t s = current_stream();
    const size_t e3 = (size_t)edge * edge * edge, ft = b->fp32 ? 4 : 8;
    char *p = b->pool.base + off;
    BoxDev &r = b->hrec[id];
    std::memset(&r, 0, sizeof r);
    r.e = edge;
    for (int d = 0; d < 3; d++) {
        r.dx[d] = dx;
        r.dxi[d] = 1.0E0 / dx;
    }
    r.cfl = cfl;
    r.eta = eta;
    r.alam = alam;
    r.U = (double *)p;
    p += align_up(5 * e3 * 8);
    r.S1 = (double *)p;
    p += align_up(5 * e3 * 8);
    r.S2 = (double *)p;
    p += align_up(5 * e3 * 8);
    for (int q = 0; q < 3; q++) {
        r.Q[q] = p;
        p += align_up(5 * e3 * ft);
    }
    r.D = p;
    p += align_up(4 * e3 * ft);
    r.WZ = p;
    int rc = CNS_OK;
    if (hipMemcpyAsync(b->d_boxes + id, &r, sizeof r, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(r.U, u0, 5 * e3 * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        launch_batch_init(b->d_boxes, id, edge, b->fp32, s) != hipSuccess ||
        launch_batch_finalize(b->d_boxes, b->d_iota + id, 1, 0, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = fail(CNS_EHIP, "initialising box %d failed: %s", id, hipGetErrorString(hipGetLastError()));
    if (rc) {
        b->pool.release(off, bytes);
        return rc;
    }
    BoxHost &h = b->hb[id];
    h.live = true;
    h.e = edge;
    h.off = off;
    h.bytes = bytes;
    b->dirty = true;
    *box_id = id;
    return CNS_OK;
}

int cns_batch_remove(cns_batch *b, int box_id)
{
    int rc = check_id(b, box_id);
    if (rc) return rc;
    // the box's memory may still be in use by queued steps
    CNS_HIPCHK(hipStreamSynchronize(current_stream()));
    BoxHost &h = b->hb[box_id];
    b->pool.release(h.off, h.bytes);
    h.live = false;
    b->dirty = true;
    return CNS_OK;
}

int cns_batch_read(cns_batch *b, int box_id, double *u)
{
    int rc = check_id(b, box_id);
    if (rc) return rc;
    if (!u) return fail(CNS_EINVAL, "NULL destination");
    const BoxHost &h = b->hb[box_id];
    const size_t e3 = (size_t)h.e * h.e * h.e;
    CNS_HIPCHK(hipMemcpyAsync(u, b->hrec[box_id].U, 5 * e3 * sizeof(double), hipMemcpyDeviceToDevice,
                              current_stream()));
    return CNS_OK;
}

int cns_batch_step(cns_batch *b, int nsteps)
{
    if (!b) return fail(CNS_EINVAL, "NULL engine");
    if (nsteps < 0) return fail(CNS_EINVAL, "negative step count");
    hipStream_t s = current_stream();
    if (b->dirty) {
        int rc = rebuild(b, s);
        if (rc) return rc;
    }
    if (!b->exec) return CNS_OK;
    for (int i = 0; i < nsteps; i++) CNS_HIPCHK(hipGraphLaunch(b->exec, s));
    return CNS_OK;
}

int cns_batch_set_target(cns_batch *b, int box_id, long long target)
{
    int rc = check_id(b, box_id);
    if (rc) return rc;
    if (target < 0) return fail(CNS_EINVAL, "negative step target");
    hipStream_t s = current_stream();
    b->hrec[box_id].target = target;
    CNS_HIPCHK(hipMemcpyAsync(&b->d_boxes[box_id].target, &b->hrec[box_id].target, sizeof(long long),
                              hipMemcpyHostToDevice, s));
    CNS_HIPCHK(hipStreamSynchronize(s));  // the staging record may change before the copy runs
    return CNS_OK;
}

int cns_batch_query(cns_batch *b, int n, const int *ids, long long *steps, double *dt, double *courno,
                    int *err)
{
    if (!b || (n > 0 && !ids)) return fail(CNS_EINVAL, "NULL argument");
    int hi = -1;
    for (int i = 0; i < n; i++) {
        int rc = check_id(b, ids[i]);
        if (rc) return rc;
        hi = std::max(hi, ids[i]);
    }
    if (hi < 0) return CNS_OK;
    std::vector<BoxDev> cur(hi + 1);
    hipStream_t s = current_stream();
    CNS_HIPCHK(hipMemcpyAsync(cur.data(), b->d_boxes, (hi + 1) * sizeof(BoxDev), hipMemcpyDeviceToHost, s));
    CNS_HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) {
        const BoxDev &r = cur[ids[i]];
        if (steps) steps[i] = r.steps;
        if (dt) dt[i] = r.dt;
        if (courno) courno[i] = r.courno;
        if (err) err[i] = r.err;
    }
    return CNS_OK;
}

int cns_batch_set_tile(cns_batch *b, int edge, int tile_width, int kz)
{
    if (!b) return fail(CNS_EINVAL, "NULL engine");
    if (width_class(tile_width) < 0) return fail(CNS_EINVAL, "tile width %d not 16, 32 or 64", tile_width);
    if (edge < 1 || kz < 1) return fail(CNS_EINVAL, "bad edge or kz");
    b->tune[edge] = {tile_width, kz};
    b->dirty = true;
    return CNS_OK;
}

int cns_batch_get_tile(cns_batch *b, int edge, int *tile_width, int *kz)
{
    if (!b || !tile_width || !kz) return fail(CNS_EINVAL, "NULL argument");
    const auto t = b->tile_of(edge);
    *tile_width = t.first;
    *kz = std::min(t.second, std::max(edge, 1));
    return CNS_OK;
}

int cns_batch_stats(cns_batch *b, size_t *used_bytes, int *boxes, int *tiles)
{
    if (!b) return fail(CNS_EINVAL, "NULL engine");
    if (b->dirty) {
        int rc = rebuild(b, current_stream());
        if (rc) return rc;
    }
    if (used_bytes) *used_bytes = b->pool.used;
    if (boxes) *boxes = b->nresident;
    if (tiles) *tiles = b->ntiles;
    return CNS_OK;
}

}  // extern "C"
