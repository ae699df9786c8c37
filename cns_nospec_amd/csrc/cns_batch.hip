// cns_batch.hip -- kernels of the batched engine: many independent periodic boxes
// of uneven edges advanced together, a few launches per RK stage.
//
// Per stage and tile-width class, three kernels walk the tile work list (tiles of
// every active box, one workgroup per tile), without ghost storage (halos are
// loaded at periodically wrapped addresses): the wz field of the stage state,
// diffterm, and hypterm with the SSP-RK3 update of each cell in FP64 and ctoprim of
// the updated cell, whose primitives the next stage reads (so no separate ctoprim
// pass exists after start-up).  In stage 3 the last also folds the Courant number
// of U^{n+1} into the box's accumulator; a small finalize kernel turns that into
// the next step's dt.  A step is captured once into a hipGraph.
//
// Flux arithmetic type FT: double reproduces the plain path (and the reference)
// bit for bit; float is the FP32-flux mode, in which only the stencil arithmetic
// of hypterm and diffterm runs in FP32 -- U, the RK update, ctoprim's p/T recovery
// and the dt reduction stay FP64.  In FP32 mode the pressure is carried as
// p - pref (pref = the box's initial pressure at cell 0) so that the pressure-
// gradient differences keep FP32's relative precision on the perturbation
// instead of on the O(1) mean; the energy flux adds pref back.
#include "cns_batch_dev.h"
#include "cns_stencil_ops.h"

namespace cns {

namespace {

constexpr int kLdsCells = 864;  // max (TY+8)*(TX+8) over TX in {16, 32, 64}

// Pointers read from a box record are generic; casting them to the global
// address space lets the compiler address through a scalar base (global_load
// saddr) instead of a per-lane 64-bit flat address.
#define GAS __attribute__((address_space(1)))

// Global array indexed by an unsigned 32-bit cell offset, turned into a 32-bit
// byte offset from a uniform base, so that each access is one scalar-base
// global_load/store with a 32-bit lane offset.  An index must stay inside ONE
// component (< e^3 <= 2^27 cells, so index * 8 < 2^30): the component base is
// formed by operator+, in 64-bit pointer arithmetic, because a box's five
// components together pass 4 GiB for edges >= 476.
template <typename T>
struct GArr {
    T GAS *p;
    __device__ __forceinline__ T GAS &operator[](unsigned i) const
    {
        return *(T GAS *)((char GAS *)p + i * (unsigned)sizeof(T));
    }
    __device__ __forceinline__ GArr operator+(size_t n) const { return GArr{p + n}; }
};
template <typename T>
__device__ __forceinline__ GArr<T> garr(void *p)
{
    return GArr<T>{(T GAS *)p};
}

// Keeps the scheduler from hoisting the loads of later groups of a long
// straight-line block above earlier ones (bounds the live registers).
#define FENCE() __builtin_amdgcn_sched_barrier(0)

__device__ __forceinline__ int wrapc(int a, int e) { return a < 0 ? a + e : (a >= e ? a - e : a); }

// ctoprim of one cell, cpu/ctoprim.cpp:31-40.
__device__ __forceinline__ void prim_cell(double r, double mx, double my, double mz, double E,
                                          double &u, double &v, double &w, double &p, double &T)
{
    double eint;
    ctoprim_cell(r, mx, my, mz, E, u, v, w, eint, p);
    T = eint / CV;
}

// Courant number of one cell, cpu/ctoprim.cpp:52-59, as the bits of its max.
__device__ __forceinline__ unsigned long long cour_bits(double r, double u, double v, double w,
                                                        double p, const double dx[3])
{
    const double c = sqrt(GAMMA * p / r);
    unsigned long long b = courno_bits((c + fabs(u)) / dx[0]);
    const unsigned long long by = courno_bits((c + fabs(v)) / dx[1]);
    const unsigned long long bz = courno_bits((c + fabs(w)) / dx[2]);
    b = b > by ? b : by;
    return b > bz ? b : bz;
}

__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v)
{
    __shared__ unsigned long long red[kBatchThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = v > o ? v : o;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long m = red[0];
#pragma unroll
    for (int w = 1; w < kBatchThreads / 64; w++) m = m > red[w] ? m : red[w];
    return m;
}

// ------------------------------------------------------------------------ init
// Primitives of U^0 into Q[0], its Courant maximum into acc, and pref.
template <typename FT>
__global__ void __launch_bounds__(kBatchThreads)
batch_init_kernel(BoxDev *__restrict__ boxes, int b)
{
    BoxDev &B = boxes[b];
    const int e = B.e;
    const int64_t e3 = (int64_t)e * e * e;
    const double *U = B.U;
    FT *q = (FT *)B.Q[0];
    double pref = 0.0;
    if (sizeof(FT) == 4) {
        double u0, v0, w0, T0;
        prim_cell(U[0], U[e3], U[2 * e3], U[3 * e3], U[4 * e3], u0, v0, w0, pref, T0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) B.pref = pref;
    const double dx[3] = {B.dx[0], B.dx[1], B.dx[2]};
    unsigned long long cmax = 0ull;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < e3;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double r = U[c];
        double u, v, w, p, T;
        prim_cell(r, U[c + e3], U[c + 2 * e3], U[c + 3 * e3], U[c + 4 * e3], u, v, w, p, T);
        q[c] = (FT)u;
        q[c + e3] = (FT)v;
        q[c + 2 * e3] = (FT)w;
        q[c + 3 * e3] = (FT)(p - pref);
        q[c + 4 * e3] = (FT)T;
        const unsigned long long cb = cour_bits(r, u, v, w, p, dx);
        cmax = cmax > cb ? cmax : cb;
    }
    const unsigned long long m = block_max_u64(cmax);
    if (threadIdx.x == 0) atomicMax(&B.acc, m);
}

// -------------------------------------------------------------------- finalize
// courno = max(acc, 1e-50), dt = cfl/courno (cpu/advance.cpp:52-58); flag a
// non-finite Courant number; clear the accumulator for the next stage 3.
__global__ void batch_finalize_kernel(BoxDev *__restrict__ boxes, const int *__restrict__ ids,
                                      int n, int count_step)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    BoxDev &B = boxes[ids[t]];
    if (box_stopped(B)) return;
    const double c = __longlong_as_double((long long)B.acc);
    double courno = c > 1.0E-50 ? c : 1.0E-50;
    if (!(c == c)) courno = c;
    B.courno = courno;
    B.dt = B.cfl / courno;
    if (!(courno <= 1.7976931348623157e308)) B.err = 1;
    B.acc = 0ull;
    if (count_step) B.steps += 1;
}

// ----------------------------------------------------------------------- tiles
// A tile is a TX x TY column block of one box (TX = 2^LTX, TX*TY = 256 threads)
// marching planes z0 .. z0+nz-1.  Both tile bodies are templates on the width so
// that every LDS neighbour offset is a compile-time immediate.  LDS cell (r, c)
// holds box cell (x0+c-4, y0+r-4) wrapped periodically; the lane (tx, ty) owns
// (r, c) = (ty+4, tx+4).  Lanes whose cell lies past the box edge (partial tiles)
// still stage and derive values (their wrapped cells are halo for others) but
// write nothing.
struct TileGeo {
    int e;
    unsigned e2, e3, own;  // cell offsets within one component fit 32 bits (edge <= 512)
    int xo, yo;
    bool active;
};

template <int LTX>
__device__ __forceinline__ TileGeo tile_geo(int e, int x0, int y0)
{
    constexpr int TXr = 1 << LTX;
    TileGeo t;
    t.e = e;
    t.e2 = e * e;
    t.e3 = t.e2 * e;
    const int tx = threadIdx.x & (TXr - 1), ty = threadIdx.x >> LTX;
    t.xo = (x0 + tx) % e;
    t.yo = (y0 + ty) % e;
    t.active = (x0 + tx < e) && (y0 + ty < e);
    t.own = t.yo * e + t.xo;
    return t;
}

// Halo cells of an LDS plane: 8*TY x-halo cells (rows 4..TY+3, cols 0..3 and
// TX+4..TX+7), 8*TX y-halo cells (rows 0..3 and TY+4..TY+7, cols 4..TX+3), 64
// corners.  Each also gives the wrapped in-plane offset gy*e + gx.
template <int LTX>
__device__ __forceinline__ void xhalo(int h, int x0, int y0, int e, int &l, int &ho)
{
    constexpr int TXr = 1 << LTX, P = TXr + 8;
    const int r = 4 + (h >> 3), cc = h & 7, c = cc < 4 ? cc : TXr + cc;
    l = r * P + c;
    ho = ((y0 + r - 4) % e) * e + wrapc((x0 + c - 4) % e, e);
}
template <int LTX>
__device__ __forceinline__ void yhalo(int h, int x0, int y0, int e, int &l, int &ho)
{
    constexpr int TXr = 1 << LTX, TYr = kBatchThreads >> LTX, P = TXr + 8;
    const int rr = h >> LTX, c = 4 + (h & (TXr - 1)), r = rr < 4 ? rr : TYr + rr;
    l = r * P + c;
    ho = wrapc((y0 + r - 4) % e, e) * e + (x0 + c - 4) % e;
}
template <int LTX>
__device__ __forceinline__ void corner(int h, int x0, int y0, int e, int &l, int &ho)
{
    constexpr int TXr = 1 << LTX, TYr = kBatchThreads >> LTX, P = TXr + 8;
    const int rr = h >> 3, cc = h & 7;
    const int r = rr < 4 ? rr : TYr + rr, c = cc < 4 ? cc : TXr + cc;
    l = r * P + c;
    ho = wrapc((y0 + r - 4) % e, e) * e + wrapc((x0 + c - 4) % e, e);
}

// ------------------------------------------------------------------ diffterm tile
// cpu/diffterm.cpp:22-320 on the stage's primitives; D (mx, my, mz, E components;
// D_rho = 0 is implicit) to the box's D buffer in the flux type.  Same three LDS
// phases per plane as the plain tiled diffterm.
template <typename FT, int LTX>
__device__ __forceinline__ void diff_tile(BoxDev *__restrict__ boxes, int b, int x0, int y0, int z0,
                                          int nz, int stage, FT *lds)
{
    using KK = K<FT>;
    constexpr int TXr = 1 << LTX, TYr = kBatchThreads >> LTX, P = TXr + 8;
    FT *const s_u = lds, *const s_v = lds + kLdsCells, *const s_w = lds + 2 * kLdsCells;
    FT *const s_T = lds + 3 * kLdsCells, *const s_ux = lds + 4 * kLdsCells;
    FT *const s_vy = lds + 5 * kLdsCells, *const s_wz = lds + 6 * kLdsCells;
    FT *const s_lu = lds + 7 * kLdsCells, *const s_lv = lds + 8 * kLdsCells;
    const int tid = threadIdx.x;
    const BoxDev &B = boxes[b];
    const TileGeo t = tile_geo<LTX>(B.e, x0, y0);
    const int e = t.e;
    const int e2 = t.e2, e3 = t.e3, own = t.own;
    const int me = ((tid >> LTX) + 4) * P + (tid & (TXr - 1)) + 4;
    const GArr<FT> qi = garr<FT>(B.Q[stage - 1]);
    const GArr<FT> q_u = qi, q_v = qi + e3, q_w = qi + 2 * e3, q_t = qi + 4 * e3;
    const GArr<FT> d_mx = garr<FT>(B.D), d_my = d_mx + e3, d_mz = d_mx + 2 * e3, d_E = d_mx + 3 * e3;
    const FT dxi0 = (FT)B.dxi[0], dxi1 = (FT)B.dxi[1], dxi2 = (FT)B.dxi[2];
    const FT dxi0s = (FT)(B.dxi[0] * B.dxi[0]), dxi1s = (FT)(B.dxi[1] * B.dxi[1]),
             dxi2s = (FT)(B.dxi[2] * B.dxi[2]);
    const FT eta = (FT)B.eta, alam = (FT)B.alam;

    FT wu[9], wv[9], ww[9], wt[9], wux[9], wvy[9];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int o = wrapc(z0 - 4 + m, e) * e2 + own;
        wu[m] = q_u[o]; wv[m] = q_v[o]; ww[m] = q_w[o]; wt[m] = q_t[o];
    }
#define LUX(mm) s_lu[me + (mm)]
#define LVY(mm) s_lv[me + (mm) * P]
    // ux, vy of the first 8 window planes (rolled: its barriers would otherwise
    // keep every plane's loads live)
#pragma unroll 1
    for (int m = 0; m < 8; m++) {
        const int pl = wrapc(z0 - 4 + m, e) * e2;
        __syncthreads();
        s_lu[me] = q_u[pl + own];
        s_lv[me] = q_v[pl + own];
        for (int h = tid; h < 8 * TYr; h += kBatchThreads) {
            int l, ho;
            xhalo<LTX>(h, x0, y0, e, l, ho);
            s_lu[l] = q_u[pl + ho];
        }
        for (int h = tid; h < 8 * TXr; h += kBatchThreads) {
            int l, ho;
            yhalo<LTX>(h, x0, y0, e, l, ho);
            s_lv[l] = q_v[pl + ho];
        }
        __syncthreads();
        const FT uxm = KD1(LUX) * dxi0, vym = KD1(LVY) * dxi1;
#pragma unroll
        for (int mm = 0; mm < 7; mm++) {
            wux[mm] = wux[mm + 1];
            wvy[mm] = wvy[mm + 1];
        }
        wux[7] = uxm;
        wvy[7] = vym;
    }

    FT nu, nv, nw, nt;  // prefetched own values of the next leading plane
    {
        const int o = wrapc(z0 + 4, e) * e2 + own;
        nu = q_u[o]; nv = q_v[o]; nw = q_w[o]; nt = q_t[o];
    }
    for (int k = z0; k < z0 + nz; k++) {
        const int cp = k * e2, lp = wrapc(k + 4, e) * e2;
        wu[8] = nu; wv[8] = nv; ww[8] = nw; wt[8] = nt;
        // phase A: lead-plane u (x halo), v (y halo); centre u, v (square), w, T (cross)
        __syncthreads();
        s_lu[me] = wu[8]; s_lv[me] = wv[8];
        s_u[me] = wu[4]; s_v[me] = wv[4]; s_w[me] = ww[4]; s_T[me] = wt[4];
        for (int h = tid; h < 8 * TYr; h += kBatchThreads) {
            int l, ho;
            xhalo<LTX>(h, x0, y0, e, l, ho);
            const int o = cp + ho;
            s_u[l] = q_u[o]; s_v[l] = q_v[o]; s_w[l] = q_w[o]; s_T[l] = q_t[o];
            s_lu[l] = q_u[lp + ho];
        }
        for (int h = tid; h < 8 * TXr; h += kBatchThreads) {
            int l, ho;
            yhalo<LTX>(h, x0, y0, e, l, ho);
            const int o = cp + ho;
            s_u[l] = q_u[o]; s_v[l] = q_v[o]; s_w[l] = q_w[o]; s_T[l] = q_t[o];
            s_lv[l] = q_v[lp + ho];
        }
        if (tid < 64) {
            int l, ho;
            corner<LTX>(tid, x0, y0, e, l, ho);
            s_u[l] = q_u[cp + ho];
            s_v[l] = q_v[cp + ho];
        }
        __syncthreads();
        // phase B: lead ux, vy into the windows; centre ux (y halo), vy (x halo), wz (cross)
        wux[8] = KD1(LUX) * dxi0;
        wvy[8] = KD1(LVY) * dxi1;
        s_ux[me] = wux[4];
        s_vy[me] = wvy[4];
#define WWZ(mm) ww[4 + (mm)]
        s_wz[me] = KD1(WWZ) * dxi2;
#undef WWZ
        for (int h = tid; h < 8 * TXr; h += kBatchThreads) {
            int l, ho;
            yhalo<LTX>(h, x0, y0, e, l, ho);
#define SUX(mm) s_u[l + (mm)]
            s_ux[l] = KD1(SUX) * dxi0;
#undef SUX
        }
        for (int h = tid; h < 8 * TYr; h += kBatchThreads) {
            int l, ho;
            xhalo<LTX>(h, x0, y0, e, l, ho);
#define SVY(mm) s_v[l + (mm) * P]
            s_vy[l] = KD1(SVY) * dxi1;
#undef SVY
        }
        for (int h = tid; h < 8 * TYr + 8 * TXr; h += kBatchThreads) {
            int l, ho;
            if (h < 8 * TYr) xhalo<LTX>(h, x0, y0, e, l, ho);
            else yhalo<LTX>(h - 8 * TYr, x0, y0, e, l, ho);
#define GWZ(mm) q_w[wrapc(k + (mm), e) * e2 + ho]
            s_wz[l] = KD1(GWZ) * dxi2;
#undef GWZ
        }
        __syncthreads();
        {   // next leading plane; its latency hides behind phase C
            const int o = wrapc(k + 5, e) * e2 + own;
            nu = q_u[o]; nv = q_v[o]; nw = q_w[o]; nt = q_t[o];
        }
        // phase C (cpu/diffterm.cpp:155-311)
        if (t.active) {
#define CUx(m) s_u[me + (m)]
#define CUy(m) s_u[me + (m) * P]
#define CUz(m) wu[4 + (m)]
#define CVx(m) s_v[me + (m)]
#define CVy(m) s_v[me + (m) * P]
#define CVz(m) wv[4 + (m)]
#define CWx(m) s_w[me + (m)]
#define CWy(m) s_w[me + (m) * P]
#define CWz(m) ww[4 + (m)]
#define CTx(m) s_T[me + (m)]
#define CTy(m) s_T[me + (m) * P]
#define CTz(m) wt[4 + (m)]
#define VYx(m) s_vy[me + (m)]
#define WZx(m) s_wz[me + (m)]
#define UXy(m) s_ux[me + (m) * P]
#define WZy(m) s_wz[me + (m) * P]
#define UXz(m) wux[4 + (m)]
#define VYz(m) wvy[4 + (m)]
            const int zo = cp + own;
            const FT uxx = KD2(CUx) * dxi0s, uyy = KD2(CUy) * dxi1s, uzz = KD2(CUz) * dxi2s;
            const FT vyx = KD1(VYx) * dxi0, wzx = KD1(WZx) * dxi0;
            const FT dmx = eta * (KK::FourThirds * uxx + uyy + uzz + KK::OneThird * (vyx + wzx));
            FENCE();
            const FT vxx = KD2(CVx) * dxi0s, vyy = KD2(CVy) * dxi1s, vzz = KD2(CVz) * dxi2s;
            const FT uxy = KD1(UXy) * dxi1, wzy = KD1(WZy) * dxi1;
            const FT dmy = eta * (vxx + KK::FourThirds * vyy + vzz + KK::OneThird * (uxy + wzy));
            FENCE();
            const FT wxx = KD2(CWx) * dxi0s, wyy = KD2(CWy) * dxi1s, wzz = KD2(CWz) * dxi2s;
            const FT uxz = KD1(UXz) * dxi2, vyz = KD1(VYz) * dxi2;
            const FT dmz = eta * (wxx + wyy + KK::FourThirds * wzz + KK::OneThird * (uxz + vyz));
            FENCE();
            const FT txx = KD2(CTx) * dxi0s, tyy = KD2(CTy) * dxi1s, tzz = KD2(CTz) * dxi2s;
            FENCE();
            const FT ux = wux[4], vy = wvy[4], wz = s_wz[me];
            const FT uy = KD1(CUy) * dxi1, uz = KD1(CUz) * dxi2;
            const FT vx = KD1(CVx) * dxi0, vz = KD1(CVz) * dxi2;
            const FT wx = KD1(CWx) * dxi0, wy = KD1(CWy) * dxi1;
            const FT divu = KK::TwoThirds * (ux + vy + wz);
            const FT tauxx = KK::Two * ux - divu, tauyy = KK::Two * vy - divu;
            const FT tauzz = KK::Two * wz - divu;
            const FT tauxy = uy + vx, tauxz = uz + wx, tauyz = vz + wy;
            FT mech = tauxx * ux + tauyy * vy + tauzz * wz + (tauxy * tauxy) + (tauxz * tauxz)
                      + (tauyz * tauyz);
            mech = eta * mech + dmx * wu[4] + dmy * wv[4] + dmz * ww[4];
            d_mx[zo] = dmx;
            d_my[zo] = dmy;
            d_mz[zo] = dmz;
            d_E[zo] = alam * (txx + tyy + tzz) + mech;
#undef CUx
#undef CUy
#undef CUz
#undef CVx
#undef CVy
#undef CVz
#undef CWx
#undef CWy
#undef CWz
#undef CTx
#undef CTy
#undef CTz
#undef VYx
#undef WZx
#undef UXy
#undef WZy
#undef UXz
#undef VYz
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wu[m] = wu[m + 1]; wv[m] = wv[m + 1]; ww[m] = ww[m + 1]; wt[m] = wt[m + 1];
            wux[m] = wux[m + 1]; wvy[m] = wvy[m + 1];
        }
    }
#undef LUX
#undef LVY
}

// ------------------------------------------------------------- staged diffterm tile
// The same arithmetic as diff_tile, with the plane data staged by direct-to-LDS loads:
// each lane moves 4 bytes from its own (periodically wrapped) address straight into a
// lane-linear LDS slot, no registers involved.  Per plane k the tile needs the centre
// plane's u, v, w, T and wz (a field written by batch_wz_kernel, so no z neighbours of
// halo columns are gathered) as full (TY+8) x (TX+8) tiles, and the leading plane
// k+4's u on the tile rows (for the ux window) and v on the tile columns (vy window).
// They go into one of two buffers; the loads of plane k+1 are issued right after the
// barrier that opens plane k, so the compute of plane k hides them.  Two barriers per
// plane: the opening one (plane k landed, plane k-1 fully consumed) and one between
// the halo derivatives of phase B and the outputs of phase C.
typedef __attribute__((address_space(3))) void lds_void;

template <typename FT, int LTX>
struct DmaGeo {
    static constexpr int TX = 1 << LTX, TY = kBatchThreads >> LTX, P = TX + 8, R = TY + 8;
    static constexpr int EPW = 256 / (int)sizeof(FT);  // elements one wave moves (4 B per lane)
    static constexpr int NW = kBatchThreads / 64;
    // regions of one buffer: full tiles (centre u, v, w, T, wz), leading u rows, leading v columns
    static constexpr int CF = R * P, CU = TY * P, CV = R * TX;
    static constexpr int SF = (CF + EPW - 1) / EPW, SU = (CU + EPW - 1) / EPW, SV = (CV + EPW - 1) / EPW;
    static constexpr int LF = SF * EPW, LU = SU * EPW, LV = SV * EPW;  // padded lengths
    static constexpr int NCF = 5;
    static constexpr int BUF = NCF * LF + LU + LV;
    // slots per wave (wave w takes slots w, w + NW, ...)
    static constexpr int WF = (SF + NW - 1) / NW, WU = (SU + NW - 1) / NW, WV = (SV + NW - 1) / NW;
    static constexpr int UXN = R * TX, VYN = TY * P;  // s_ux: R rows x TX cols; s_vy: TY rows x P cols
    static constexpr int LDS_ELEMS = 2 * BUF + UXN + VYN;
};

__device__ __forceinline__ int pmod(int a, int e) { const int r = a % e; return r < 0 ? r + e : r; }

// In-plane offset (y*e + x, periodically wrapped) of element `el` of a region whose
// row r / column c map to box cell (x0 + c - 4 + cshift, y0 + r - 4 + rshift); elements
// past the region's end (slot padding) load cell (0, 0) into the padding.
__device__ __forceinline__ unsigned region_off(int el, int n, int pitch, int rshift, int cshift, int x0,
                                               int y0, int e)
{
    if (el >= n) return 0u;
    const int r = el / pitch, c = el - r * pitch;
    return (unsigned)(pmod(y0 + r - 4 + rshift, e) * e + pmod(x0 + c - 4 + cshift, e));
}

template <typename FT, int LTX>
__device__ __forceinline__ void diff_tile_dma(BoxDev *__restrict__ boxes, int b, int x0, int y0, int z0,
                                              int nz, int stage, FT *lds)
{
    using KK = K<FT>;
    using G = DmaGeo<FT, LTX>;
    constexpr int TXr = G::TX, P = G::P, EPW = G::EPW;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const BoxDev &B = boxes[b];
    const TileGeo t = tile_geo<LTX>(B.e, x0, y0);
    const int e = t.e;
    const unsigned e2 = t.e2, e3 = t.e3, own = t.own;
    const int row = tid >> LTX, col = tid & (TXr - 1);
    const int me = (row + 4) * P + col + 4;  // own cell in a full tile
    const GArr<FT> qi = garr<FT>(B.Q[stage - 1]);
    const GArr<FT> q_u = qi, q_v = qi + e3, q_w = qi + 2 * (size_t)e3, q_t = qi + 4 * (size_t)e3;
    const GArr<FT> wzf = garr<FT>(B.WZ);
    const GArr<FT> d_mx = garr<FT>(B.D), d_my = d_mx + e3, d_mz = d_mx + 2 * (size_t)e3,
                   d_E = d_mx + 3 * (size_t)e3;
    const FT dxi0 = (FT)B.dxi[0], dxi1 = (FT)B.dxi[1], dxi2 = (FT)B.dxi[2];
    const FT dxi0s = (FT)(B.dxi[0] * B.dxi[0]), dxi1s = (FT)(B.dxi[1] * B.dxi[1]),
             dxi2s = (FT)(B.dxi[2] * B.dxi[2]);
    const FT eta = (FT)B.eta, alam = (FT)B.alam;
    FT *const s_ux = lds + 2 * G::BUF, *const s_vy = s_ux + G::UXN;

    // Per-lane in-plane byte offsets of this wave's slots (the same for every plane).
    // A 4-byte lane of an 8-byte element moves its low (even lane) or high half.
    constexpr int SUB = (int)sizeof(FT) / 4;
    const int lel = lane / SUB;
    const unsigned half = (unsigned)(lane % SUB) * 4u;
    unsigned of[G::WF], ou[G::WU], ov[G::WV];
#pragma unroll
    for (int j = 0; j < G::WF; j++)
        of[j] = region_off((wave + j * G::NW) * EPW + lel, G::CF, P, 0, 0, x0, y0, e) * sizeof(FT) + half;
#pragma unroll
    for (int j = 0; j < G::WU; j++)
        ou[j] = region_off((wave + j * G::NW) * EPW + lel, G::CU, P, 4, 0, x0, y0, e) * sizeof(FT) + half;
#pragma unroll
    for (int j = 0; j < G::WV; j++)
        ov[j] = region_off((wave + j * G::NW) * EPW + lel, G::CV, TXr, 0, 4, x0, y0, e) * sizeof(FT) + half;

    const char GAS *const src[G::NCF] = {(const char GAS *)q_u.p, (const char GAS *)q_v.p,
                                         (const char GAS *)q_w.p, (const char GAS *)q_t.p,
                                         (const char GAS *)wzf.p};
    // plane k (centre fields) and its leading plane k+4 (u rows, v columns) into buffer `buf`
    auto issue = [&](int k, FT *buf) {
        const unsigned pc = (unsigned)k * e2 * (unsigned)sizeof(FT);
        const unsigned pl = (unsigned)pmod(k + 4, e) * e2 * (unsigned)sizeof(FT);
#pragma unroll
        for (int j = 0; j < G::WF; j++) {
            const int s = wave + j * G::NW;
            if (s < G::SF) {
#pragma unroll
                for (int f = 0; f < G::NCF; f++)
                    __builtin_amdgcn_global_load_lds((const void *)(src[f] + (pc + of[j])),
                                                     (lds_void *)(buf + f * G::LF + s * EPW), 4, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < G::WU; j++) {
            const int s = wave + j * G::NW;
            if (s < G::SU)
                __builtin_amdgcn_global_load_lds((const void *)(src[0] + (pl + ou[j])),
                                                 (lds_void *)(buf + G::NCF * G::LF + s * EPW), 4, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < G::WV; j++) {
            const int s = wave + j * G::NW;
            if (s < G::SV)
                __builtin_amdgcn_global_load_lds((const void *)(src[1] + (pl + ov[j])),
                                                 (lds_void *)(buf + G::NCF * G::LF + G::LU + s * EPW), 4, 0, 0);
        }
    };
    issue(z0, lds);

    // windows: plane k+m at [4+m]; the first 8 planes and their ux, vy straight from
    // global memory (the same expressions as the per-plane path)
    FT wu[9], wv[9], ww[9], wt[9], wux[9], wvy[9];
#pragma unroll 1
    for (int m = 0; m < 8; m++) {
        const unsigned pl = (unsigned)pmod(z0 - 4 + m, e) * e2;
        const unsigned rowo = pl + (unsigned)t.yo * e;
#define GUX(mm) q_u[rowo + (unsigned)wrapc(t.xo + (mm), e)]
#define GVY(mm) q_v[pl + (unsigned)wrapc(t.yo + (mm), e) * e + (unsigned)t.xo]
        const FT uxm = KD1(GUX) * dxi0, vym = KD1(GVY) * dxi1;
#undef GUX
#undef GVY
#pragma unroll
        for (int mm = 0; mm < 7; mm++) {
            wux[mm] = wux[mm + 1];
            wvy[mm] = wvy[mm + 1];
        }
        wux[7] = uxm;
        wvy[7] = vym;
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const unsigned o = (unsigned)pmod(z0 - 4 + m, e) * e2 + own;
        wu[m] = q_u[o]; wv[m] = q_v[o]; ww[m] = q_w[o]; wt[m] = q_t[o];
    }
    FT nw, nt;  // own w, T of the next leading plane
    {
        const unsigned o = (unsigned)pmod(z0 + 4, e) * e2 + own;
        nw = q_w[o];
        nt = q_t[o];
    }

    for (int k = z0; k < z0 + nz; k++) {
        FT *const bc = lds + ((k - z0) & 1) * G::BUF;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (k + 1 < z0 + nz) issue(k + 1, lds + (((k - z0) & 1) ^ 1) * G::BUF);
        const FT *s_cu = bc, *s_cv = bc + G::LF, *s_cw = bc + 2 * G::LF, *s_ct = bc + 3 * G::LF,
                 *s_wz = bc + 4 * G::LF, *s_lu = bc + 5 * G::LF, *s_lv = bc + 5 * G::LF + G::LU;
        const int mlu = row * P + col + 4, mlv = (row + 4) * TXr + col;
        wu[8] = s_lu[mlu];
        wv[8] = s_lv[mlv];
        ww[8] = nw;
        wt[8] = nt;
        if (k + 1 < z0 + nz) {
            const unsigned o = (unsigned)pmod(k + 5, e) * e2 + own;
            nw = q_w[o];
            nt = q_t[o];
        }
        // phase B: leading-plane ux, vy into the windows; centre ux on the y-halo rows,
        // vy on the x-halo columns (cpu/diffterm.cpp:77-81, :110-114)
#define LUX(mm) s_lu[mlu + (mm)]
#define LVY(mm) s_lv[mlv + (mm) * TXr]
        wux[8] = KD1(LUX) * dxi0;
        wvy[8] = KD1(LVY) * dxi1;
#undef LUX
#undef LVY
        s_ux[(row + 4) * TXr + col] = wux[4];
        for (int h = tid; h < 8 * TXr; h += kBatchThreads) {
            const int rr = h >> LTX, r = rr < 4 ? rr : rr + G::TY, c = h & (TXr - 1);
#define SUX(mm) s_cu[r * P + c + 4 + (mm)]
            s_ux[r * TXr + c] = KD1(SUX) * dxi0;
#undef SUX
        }
        s_vy[row * P + col + 4] = wvy[4];
        for (int h = tid; h < 8 * G::TY; h += kBatchThreads) {
            const int rr = h >> 3, cc = h & 7, c = cc < 4 ? cc : cc + TXr;
#define SVY(mm) s_cv[(rr + 4 + (mm)) * P + c]
            s_vy[rr * P + c] = KD1(SVY) * dxi1;
#undef SVY
        }
        __syncthreads();
        // phase C (cpu/diffterm.cpp:155-311)
        if (t.active) {
#define CUx(m) s_cu[me + (m)]
#define CUy(m) s_cu[me + (m) * P]
#define CUz(m) wu[4 + (m)]
#define CVx(m) s_cv[me + (m)]
#define CVy(m) s_cv[me + (m) * P]
#define CVz(m) wv[4 + (m)]
#define CWx(m) s_cw[me + (m)]
#define CWy(m) s_cw[me + (m) * P]
#define CWz(m) ww[4 + (m)]
#define CTx(m) s_ct[me + (m)]
#define CTy(m) s_ct[me + (m) * P]
#define CTz(m) wt[4 + (m)]
#define VYx(m) s_vy[row * P + col + 4 + (m)]
#define WZx(m) s_wz[me + (m)]
#define UXy(m) s_ux[(row + 4 + (m)) * TXr + col]
#define WZy(m) s_wz[me + (m) * P]
#define UXz(m) wux[4 + (m)]
#define VYz(m) wvy[4 + (m)]
            const unsigned zo = (unsigned)k * e2 + own;
            const FT uxx = KD2(CUx) * dxi0s, uyy = KD2(CUy) * dxi1s, uzz = KD2(CUz) * dxi2s;
            const FT vyx = KD1(VYx) * dxi0, wzx = KD1(WZx) * dxi0;
            const FT dmx = eta * (KK::FourThirds * uxx + uyy + uzz + KK::OneThird * (vyx + wzx));
            FENCE();
            const FT vxx = KD2(CVx) * dxi0s, vyy = KD2(CVy) * dxi1s, vzz = KD2(CVz) * dxi2s;
            const FT uxy = KD1(UXy) * dxi1, wzy = KD1(WZy) * dxi1;
            const FT dmy = eta * (vxx + KK::FourThirds * vyy + vzz + KK::OneThird * (uxy + wzy));
            FENCE();
            const FT wxx = KD2(CWx) * dxi0s, wyy = KD2(CWy) * dxi1s, wzz = KD2(CWz) * dxi2s;
            const FT uxz = KD1(UXz) * dxi2, vyz = KD1(VYz) * dxi2;
            const FT dmz = eta * (wxx + wyy + KK::FourThirds * wzz + KK::OneThird * (uxz + vyz));
            FENCE();
            const FT txx = KD2(CTx) * dxi0s, tyy = KD2(CTy) * dxi1s, tzz = KD2(CTz) * dxi2s;
            FENCE();
            const FT ux = wux[4], vy = wvy[4], wz = s_wz[me];
            const FT uy = KD1(CUy) * dxi1, uz = KD1(CUz) * dxi2;
            const FT vx = KD1(CVx) * dxi0, vz = KD1(CVz) * dxi2;
            const FT wx = KD1(CWx) * dxi0, wy = KD1(CWy) * dxi1;
            const FT divu = KK::TwoThirds * (ux + vy + wz);
            const FT tauxx = KK::Two * ux - divu, tauyy = KK::Two * vy - divu;
            const FT tauzz = KK::Two * wz - divu;
            const FT tauxy = uy + vx, tauxz = uz + wx, tauyz = vz + wy;
            FT mech = tauxx * ux + tauyy * vy + tauzz * wz + (tauxy * tauxy) + (tauxz * tauxz)
                      + (tauyz * tauyz);
            mech = eta * mech + dmx * wu[4] + dmy * wv[4] + dmz * ww[4];
            d_mx[zo] = dmx;
            d_my[zo] = dmy;
            d_mz[zo] = dmz;
            d_E[zo] = alam * (txx + tyy + tzz) + mech;
#undef CUx
#undef CUy
#undef CUz
#undef CVx
#undef CVy
#undef CVz
#undef CWx
#undef CWy
#undef CWz
#undef CTx
#undef CTy
#undef CTz
#undef VYx
#undef WZx
#undef UXy
#undef WZy
#undef UXz
#undef VYz
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wu[m] = wu[m + 1]; wv[m] = wv[m + 1]; ww[m] = ww[m + 1]; wt[m] = wt[m + 1];
            wux[m] = wux[m + 1]; wvy[m] = wvy[m + 1];
        }
    }
}

// wz = D1_z(w) * dxinv(3) of the stage's primitives (cpu/diffterm.cpp:127-151, the same
// expression as diff_tile's) for the cells of one tile's columns and z run.
template <typename FT, int LTX>
__device__ __forceinline__ void wz_tile(BoxDev *__restrict__ boxes, int b, int x0, int y0, int z0, int nz,
                                        int stage)
{
    using KK = K<FT>;
    const BoxDev &B = boxes[b];
    const TileGeo t = tile_geo<LTX>(B.e, x0, y0);
    if (!t.active) return;
    const int e = t.e;
    const unsigned e2 = t.e2, own = t.own;
    const GArr<FT> q_w = garr<FT>(B.Q[stage - 1]) + 2 * (size_t)t.e3;
    const GArr<FT> wzf = garr<FT>(B.WZ);
    const FT dxi2 = (FT)B.dxi[2];
    FT w[9];
#pragma unroll
    for (int m = 0; m < 8; m++) w[m] = q_w[(unsigned)pmod(z0 - 4 + m, e) * e2 + own];
    for (int k = z0; k < z0 + nz; k++) {
        w[8] = q_w[(unsigned)pmod(k + 4, e) * e2 + own];
#define WZW(mm) w[4 + (mm)]
        wzf[(unsigned)k * e2 + own] = KD1(WZW) * dxi2;
#undef WZW
#pragma unroll
        for (int m = 0; m < 8; m++) w[m] = w[m + 1];
    }
}

// ------------------------------------------------------------- hypterm+update tile
// cpu/hypterm.cpp:11-209 on the stage state, then per cell the SSP-RK3 update in
// FP64 with the tile's D (cpu/advance.cpp:82, :114-116, :148-150), ctoprim of the
// result into the next stage's primitives, and in stage 3 the Courant maximum.
// Returns the lane's Courant maximum (bits; stage 3 only).  The stage is a template
// parameter: stage 1 then has no base operand at all (its base is the stage state).
// KEEP (FP32 flux only): the FP64 conserved values loaded as the z window's leading
// plane stay in a 5-plane register FIFO until that plane is the centre, where the
// RK update needs them, instead of being loaded again four planes later.
template <typename FT, int LTX, int stage, bool KEEP>
__device__ __forceinline__ unsigned long long hyp_tile(BoxDev *__restrict__ boxes, int b, int x0,
                                                       int y0, int z0, int nz, FT *lds)
{
    using KK = K<FT>;
    constexpr int TXr = 1 << LTX, TYr = kBatchThreads >> LTX, P = TXr + 8;
    FT *const s_mx = lds, *const s_my = lds + kLdsCells, *const s_mz = lds + 2 * kLdsCells;
    FT *const s_E = lds + 3 * kLdsCells, *const s_p = lds + 4 * kLdsCells;
    FT *const s_u = lds + 5 * kLdsCells, *const s_v = lds + 6 * kLdsCells;
    const int tid = threadIdx.x;
    const BoxDev &B = boxes[b];
    const TileGeo t = tile_geo<LTX>(B.e, x0, y0);
    const int e = t.e;
    const int e2 = t.e2, e3 = t.e3, own = t.own;
    const int me = ((tid >> LTX) + 4) * P + (tid & (TXr - 1)) + 4;
    const double *cur = stage == 1 ? B.U : (stage == 2 ? B.S1 : B.S2);
    double *out = stage == 1 ? B.S1 : (stage == 2 ? B.S2 : B.U);
    // per-component uniform bases: every access is scalar base + 32-bit offset
    const GArr<double> c_r = garr<double>((void *)cur), c_mx = c_r + e3, c_my = c_r + 2 * e3,
                       c_mz = c_r + 3 * e3, c_E = c_r + 4 * e3;
    const GArr<double> b_0 = garr<double>(B.U), o_0 = garr<double>(out);
    GArr<double> b_c[NC], o_c[NC];
#pragma unroll
    for (int l = 0; l < NC; l++) {
        b_c[l] = b_0 + (size_t)l * e3;
        o_c[l] = o_0 + (size_t)l * e3;
    }
    const GArr<FT> qi = garr<FT>(B.Q[stage - 1]), qo = garr<FT>(B.Q[stage % 3]);
    const GArr<FT> qo_v = qo + e3, qo_w = qo + 2 * e3, qo_p = qo + 3 * e3, qo_t = qo + 4 * e3;
    const GArr<FT> din = garr<FT>(B.D);
    const GArr<FT> d_c[4] = {din, din + e3, din + 2 * e3, din + 3 * e3};
    const GArr<FT> q_u = qi, q_v = qi + e3, q_w = qi + 2 * e3, q_p = qi + 3 * e3;
    const double dt = B.dt;
    const double ra = stage == 1 ? 0.0E0 : (stage == 2 ? ThreeQuarters : OneThird);
    const double rb = stage == 1 ? 1.0E0 : (stage == 2 ? OneQuarter : TwoThirds);
    const double dxd[3] = {B.dx[0], B.dx[1], B.dx[2]};
    const double pref = B.pref;
    const FT dxi0 = (FT)B.dxi[0], dxi1 = (FT)B.dxi[1], dxi2 = (FT)B.dxi[2], prefF = (FT)pref;

    // z windows: w and p always; in the FP32 mode also mx, my, mz, E (as the flux type),
    // whose z taps would otherwise be 32 global loads per cell per plane that mostly
    // miss L2 (the FP64 mode has no registers to spare for them).
    constexpr bool CWIN = sizeof(FT) == 4;
    constexpr bool FIFO = KEEP && CWIN;
    FT ww[9], wp[9], wmx[9], wmy[9], wmz[9], wE[9];
    double fmx[5], fmy[5], fmz[5], fE[5];  // FIFO: planes k .. k+4 (FIFO only)
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int o = wrapc(z0 - 4 + m, e) * e2 + own;
        ww[m] = q_w[o]; wp[m] = q_p[o];
        if (FIFO && m >= 4) {
            fmx[m - 4] = c_mx[o]; fmy[m - 4] = c_my[o]; fmz[m - 4] = c_mz[o]; fE[m - 4] = c_E[o];
            wmx[m] = (FT)fmx[m - 4]; wmy[m] = (FT)fmy[m - 4]; wmz[m] = (FT)fmz[m - 4]; wE[m] = (FT)fE[m - 4];
        } else if (CWIN) {
            wmx[m] = (FT)c_mx[o]; wmy[m] = (FT)c_my[o]; wmz[m] = (FT)c_mz[o]; wE[m] = (FT)c_E[o];
        }
    }
    // prefetched: next leading plane (w, p; conserved into the FIFO or as the flux type)
    // and next centre (u, v, and the conserved values when there is no FIFO)
    FT nw, np, nu, nv, lmx = 0, lmy = 0, lmz = 0, lE = 0;
    double nmx = 0.0, nmy = 0.0, nmz = 0.0, nE = 0.0;
    {
        const int o = wrapc(z0 + 4, e) * e2 + own, oc = z0 * e2 + own;
        nw = q_w[o]; np = q_p[o];
        nu = q_u[oc]; nv = q_v[oc];
        if (FIFO) {
            fmx[4] = c_mx[o]; fmy[4] = c_my[o]; fmz[4] = c_mz[o]; fE[4] = c_E[o];
        } else {
            nmx = c_mx[oc]; nmy = c_my[oc]; nmz = c_mz[oc]; nE = c_E[oc];
            if (CWIN) {
                lmx = (FT)c_mx[o]; lmy = (FT)c_my[o]; lmz = (FT)c_mz[o]; lE = (FT)c_E[o];
            }
        }
    }
    unsigned long long cmax = 0ull;
    for (int k = z0; k < z0 + nz; k++) {
        const int cp = k * e2;
        const int zo = cp + own;
        ww[8] = nw; wp[8] = np;
        if (FIFO) {
            wmx[8] = (FT)fmx[4]; wmy[8] = (FT)fmy[4]; wmz[8] = (FT)fmz[4]; wE[8] = (FT)fE[4];
        } else if (CWIN) {
            wmx[8] = lmx; wmy[8] = lmy; wmz[8] = lmz; wE[8] = lE;
        }
        const double cmx = FIFO ? fmx[0] : nmx, cmy = FIFO ? fmy[0] : nmy;
        const double cmz = FIFO ? fmz[0] : nmz, cE = FIFO ? fE[0] : nE;
        if (FIFO) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                fmx[m] = fmx[m + 1]; fmy[m] = fmy[m + 1]; fmz[m] = fmz[m + 1]; fE[m] = fE[m + 1];
            }
        }
        const FT u_own = nu, v_own = nv;
        const FT mx_own = (FT)cmx, my_own = (FT)cmy, mz_own = (FT)cmz, E_own = (FT)cE;
        const double cur_c[5] = {0.0, cmx, cmy, cmz, cE};
        __syncthreads();
        s_mx[me] = mx_own; s_my[me] = my_own; s_mz[me] = mz_own; s_E[me] = E_own; s_p[me] = wp[4];
        s_u[me] = u_own; s_v[me] = v_own;
        for (int h = tid; h < 8 * TYr; h += kBatchThreads) {
            int l, ho;
            xhalo<LTX>(h, x0, y0, e, l, ho);
            const int o = cp + ho;
            s_mx[l] = (FT)c_mx[o]; s_my[l] = (FT)c_my[o]; s_mz[l] = (FT)c_mz[o]; s_E[l] = (FT)c_E[o];
            s_p[l] = q_p[o]; s_u[l] = q_u[o];
        }
        for (int h = tid; h < 8 * TXr; h += kBatchThreads) {
            int l, ho;
            yhalo<LTX>(h, x0, y0, e, l, ho);
            const int o = cp + ho;
            s_mx[l] = (FT)c_mx[o]; s_my[l] = (FT)c_my[o]; s_mz[l] = (FT)c_mz[o]; s_E[l] = (FT)c_E[o];
            s_p[l] = q_p[o]; s_v[l] = q_v[o];
        }
        __syncthreads();
        if (k + 1 < z0 + nz) {  // next iteration's operands, hidden behind the compute
            const int o = wrapc(k + 5, e) * e2 + own, oc = zo + e2;
            nw = q_w[o]; np = q_p[o];
            nu = q_u[oc]; nv = q_v[oc];
            if (FIFO) {
                fmx[4] = c_mx[o]; fmy[4] = c_my[o]; fmz[4] = c_mz[o]; fE[4] = c_E[o];
            } else {
                nmx = c_mx[oc]; nmy = c_my[oc]; nmz = c_mz[oc]; nE = c_E[oc];
                if (CWIN) {
                    lmx = (FT)c_mx[o]; lmy = (FT)c_my[o]; lmz = (FT)c_mz[o]; lE = (FT)c_E[o];
                }
            }
        }
        double rkd[4], rkb[5], rho_c = 0.0;
        if (t.active) {  // this cell's RK operands
            rho_c = c_r[zo];
#pragma unroll
            for (int l = 0; l < 4; l++) rkd[l] = (double)d_c[l][zo];
            if (stage != 1) {
#pragma unroll
                for (int l = 0; l < NC; l++) rkb[l] = b_c[l][zo];
            }
        }
        if (t.active) {
            FT F[5];
#define XMX(m) s_mx[me + (m)]
#define XMY(m) s_my[me + (m)]
#define XMZ(m) s_mz[me + (m)]
#define XE(m) s_E[me + (m)]
#define XP(m) s_p[me + (m)]
#define XPF(m) (s_p[me + (m)] + prefF)
#define XU(m) s_u[me + (m)]
            F[IRHO] = -(KD1(XMX)) * dxi0;
            F[IMX] = -(KHMOMP(XMX, XU, XP)) * dxi0;
            F[IMY] = -(KHMOM(XMY, XU)) * dxi0;
            F[IMZ] = -(KHMOM(XMZ, XU)) * dxi0;
            F[IENE] = -(KHENE(XE, XU, XPF)) * dxi0;
            FENCE();
#define YMX(m) s_mx[me + (m) * P]
#define YMY(m) s_my[me + (m) * P]
#define YMZ(m) s_mz[me + (m) * P]
#define YE(m) s_E[me + (m) * P]
#define YP(m) s_p[me + (m) * P]
#define YPF(m) (s_p[me + (m) * P] + prefF)
#define YV(m) s_v[me + (m) * P]
            F[IRHO] = F[IRHO] - (KD1(YMY)) * dxi1;
            F[IMX] = F[IMX] - (KHMOM(YMX, YV)) * dxi1;
            F[IMY] = F[IMY] - (KHMOMP(YMY, YV, YP)) * dxi1;
            F[IMZ] = F[IMZ] - (KHMOM(YMZ, YV)) * dxi1;
            F[IENE] = F[IENE] - (KHENE(YE, YV, YPF)) * dxi1;
            FENCE();
#define ZPL(m) (wrapc(k + (m), e) * e2 + own)
#define ZMX(m) (CWIN ? wmx[4 + (m)] : (FT)c_mx[ZPL(m)])
#define ZMY(m) (CWIN ? wmy[4 + (m)] : (FT)c_my[ZPL(m)])
#define ZMZ(m) (CWIN ? wmz[4 + (m)] : (FT)c_mz[ZPL(m)])
#define ZE(m) (CWIN ? wE[4 + (m)] : (FT)c_E[ZPL(m)])
#define ZW(m) ww[4 + (m)]
#define ZP(m) wp[4 + (m)]
#define ZPF(m) (wp[4 + (m)] + prefF)
            F[IRHO] = F[IRHO] - (KD1(ZMZ)) * dxi2;
            F[IMX] = F[IMX] - (KHMOM(ZMX, ZW)) * dxi2;
            F[IMY] = F[IMY] - (KHMOM(ZMY, ZW)) * dxi2;
            F[IMZ] = F[IMZ] - (KHMOMP(ZMZ, ZW, ZP)) * dxi2;
            F[IENE] = F[IENE] - (KHENE(ZE, ZW, ZPF)) * dxi2;
            FENCE();
            // stage 1: out = cur + dt*(D+F) (cpu/advance.cpp:82); stages 2 and 3:
            // out = ra*base + rb*(cur + dt*(D+F)), (ra, rb) = (3/4, 1/4) and (1/3, 2/3)
            // (:114-116, :148-150)
            double o[5];
#pragma unroll
            for (int l = 0; l < NC; l++) {
                const double dl = l == 0 ? 0.0E0 : rkd[l - 1];
                const double df = dl + (double)F[l];
                const double x = (l == 0 ? rho_c : cur_c[l]) + dt * df;
                o[l] = stage == 1 ? x : ra * rkb[l] + rb * x;
                o_c[l][zo] = o[l];
            }
            double u, v, w, p, T;
            prim_cell(o[0], o[1], o[2], o[3], o[4], u, v, w, p, T);
            qo[zo] = (FT)u;
            qo_v[zo] = (FT)v;
            qo_w[zo] = (FT)w;
            qo_p[zo] = (FT)(p - pref);
            qo_t[zo] = (FT)T;
            if (stage == 3) {
                const unsigned long long cb = cour_bits(o[0], u, v, w, p, dxd);
                cmax = cmax > cb ? cmax : cb;
            }
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            ww[m] = ww[m + 1];
            wp[m] = wp[m + 1];
            if (CWIN) {
                wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1]; wE[m] = wE[m + 1];
            }
        }
    }
    return cmax;
}
#undef XMX
#undef XMY
#undef XMZ
#undef XE
#undef XP
#undef XPF
#undef XU
#undef YMX
#undef YMY
#undef YMZ
#undef YE
#undef YP
#undef YPF
#undef YV
#undef ZPL
#undef ZMX
#undef ZMY
#undef ZMZ
#undef ZE
#undef ZW
#undef ZP
#undef ZPF

// ------------------------------------------------------- staged hypterm+update tile
// The hypterm + update of 32 x 8 tiles with each plane staged by direct-to-LDS loads
// (4 bytes per lane, wrapped addresses), double-buffered, one barrier per plane.  mx,
// my, mz, E are staged as FP64 on the tile's cross -- the own rows with their x halos
// (10 wave-loads of 32 values per field) and the eight y-halo row runs of 32 values
// (one wave-load each); corners are never loaded -- and p, v as full tiles, u on the
// own rows, in the flux type (an FP64 value is two 4-byte lanes).  In the FP32-flux
// mode the flux arithmetic converts each FP64 value to FP32 where it reads it (the same
// bits as hyp_tile, which converts on load); in the FP64 mode every value is used as
// loaded.  The cell's own FP64 values for the RK update come from the staged tile, so
// nothing is loaded twice by the lane that owns it.
struct HypDmaBase {
    static constexpr int TX = 32, TY = 8, P = 40, R = 16, TILE = R * P;
    static constexpr int NW = kBatchThreads / 64;
    static constexpr int XS0 = 4 * P, XSN = TY * P;         // own-row strip: elements 160..479
    static constexpr int SX64 = XSN / 32, SY64 = 8;          // FP64 position slots: strip, halo rows
    static constexpr int S64 = SX64 + SY64;                  // 18
    static constexpr int W64 = (S64 + NW - 1) / NW;
    static constexpr int B64 = 4 * TILE * 8;                 // bytes of the 4 FP64 tiles
};
template <typename FT>
struct HypDmaGeo : HypDmaBase {
    static constexpr int FB = (int)sizeof(FT);
    static constexpr int EPS = 256 / FB;                     // flux-type elements per wave-load
    static constexpr int SF = TILE / EPS, SU = XSN / EPS;    // flux-type slots: full tile, strip
    static constexpr int WF = (SF + NW - 1) / NW, WU = (SU + NW - 1) / NW;
    // buffer layout in bytes: 4 FP64 full tiles, then p and v full tiles, then the u strip
    static constexpr int BP = B64, BV = BP + TILE * FB, BU = BV + TILE * FB;
    static constexpr int BUF = BU + XSN * FB;                // 26880 (FP32) / 33280 (FP64) bytes
    static_assert(TILE % EPS == 0 && XSN % EPS == 0, "whole wave-loads per flux-type region");
    static_assert(XSN % 32 == 0 && BUF % 8 == 0, "whole FP64 wave-loads; buffers of whole doubles");
};

// Element of FP64 position slot s (lane pair L/2) in the full-tile layout.
__device__ __forceinline__ int hyp64_elem(int s, int l2)
{
    using G = HypDmaBase;
    if (s < G::SX64) return G::XS0 + 32 * s + l2;
    const int hr = s - G::SX64, r = hr < 4 ? hr : hr + G::TY;  // rows 0..3, 12..15
    return r * G::P + 4 + l2;
}

template <typename FT, int stage>
__device__ __forceinline__ unsigned long long hyp_tile_dma(BoxDev *__restrict__ boxes, int b, int x0, int y0,
                                                           int z0, int nz, char *lds)
{
    using KK = K<FT>;
    using G = HypDmaGeo<FT>;
    constexpr int LTX = 5, TXr = G::TX, P = G::P;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const BoxDev &B = boxes[b];
    const TileGeo t = tile_geo<LTX>(B.e, x0, y0);
    const int e = t.e;
    const unsigned e2 = t.e2, e3 = t.e3, own = t.own;
    const int row = tid >> LTX, col = tid & (TXr - 1);
    const int me = (row + 4) * P + col + 4;  // own cell in the full-tile layout
    const int mu = row * P + col + 4;        // own cell in the u strip
    const double *cur = stage == 1 ? B.U : (stage == 2 ? B.S1 : B.S2);
    double *out = stage == 1 ? B.S1 : (stage == 2 ? B.S2 : B.U);
    const GArr<double> c_r = garr<double>((void *)cur), c_mx = c_r + e3, c_my = c_r + 2 * (size_t)e3,
                       c_mz = c_r + 3 * (size_t)e3, c_E = c_r + 4 * (size_t)e3;
    const GArr<double> b_0 = garr<double>(B.U), o_0 = garr<double>(out);
    GArr<double> b_c[NC], o_c[NC];
#pragma unroll
    for (int l = 0; l < NC; l++) {
        b_c[l] = b_0 + (size_t)l * e3;
        o_c[l] = o_0 + (size_t)l * e3;
    }
    const GArr<FT> qi = garr<FT>(B.Q[stage - 1]), qo = garr<FT>(B.Q[stage % 3]);
    const GArr<FT> qo_v = qo + e3, qo_w = qo + 2 * (size_t)e3, qo_p = qo + 3 * (size_t)e3,
                   qo_t = qo + 4 * (size_t)e3;
    const GArr<FT> din = garr<FT>(B.D);
    const GArr<FT> d_c[4] = {din, din + e3, din + 2 * (size_t)e3, din + 3 * (size_t)e3};
    const GArr<FT> q_u = qi, q_v = qi + e3, q_w = qi + 2 * (size_t)e3, q_p = qi + 3 * (size_t)e3;
    const double dt = B.dt;
    const double ra = stage == 1 ? 0.0E0 : (stage == 2 ? ThreeQuarters : OneThird);
    const double rb = stage == 1 ? 1.0E0 : (stage == 2 ? OneQuarter : TwoThirds);
    const double dxd[3] = {B.dx[0], B.dx[1], B.dx[2]};
    const double pref = B.pref;
    const FT dxi0 = (FT)B.dxi[0], dxi1 = (FT)B.dxi[1], dxi2 = (FT)B.dxi[2], prefF = (FT)pref;

    // per-lane in-plane byte offsets of this wave's slots (the same for every plane)
    unsigned o64[G::W64], of[G::WF], ou[G::WU];
#pragma unroll
    for (int j = 0; j < G::W64; j++) {
        const int s = wave + j * G::NW;
        o64[j] = s < G::S64 ? region_off(hyp64_elem(s, lane >> 1), G::TILE, P, 0, 0, x0, y0, e) * 8u +
                                  (unsigned)(lane & 1) * 4u
                            : 0u;
    }
    // flux-type slots: element lane / (FB / 4) of the slot, and for FP64 the lane's half
    constexpr int LPE = G::FB / 4;
    const unsigned half = (unsigned)(lane & (LPE - 1)) * 4u;
#pragma unroll
    for (int j = 0; j < G::WF; j++)
        of[j] = region_off((wave + j * G::NW) * G::EPS + lane / LPE, G::TILE, P, 0, 0, x0, y0, e) * G::FB + half;
#pragma unroll
    for (int j = 0; j < G::WU; j++)
        ou[j] = region_off((wave + j * G::NW) * G::EPS + lane / LPE, G::XSN, P, 4, 0, x0, y0, e) * G::FB + half;
    const char GAS *const src64[4] = {(const char GAS *)c_mx.p, (const char GAS *)c_my.p,
                                      (const char GAS *)c_mz.p, (const char GAS *)c_E.p};
    const char GAS *const srcp = (const char GAS *)q_p.p, *const srcv = (const char GAS *)q_v.p,
                            *const srcu = (const char GAS *)q_u.p;
    auto issue = [&](int k, char *buf) {
        const unsigned p8 = (unsigned)k * e2 * 8u, p4 = (unsigned)k * e2 * (unsigned)G::FB;
#pragma unroll
        for (int j = 0; j < G::W64; j++) {
            const int s = wave + j * G::NW;
            if (s < G::S64) {
                const int el0 = s < G::SX64 ? G::XS0 + 32 * s : hyp64_elem(s, 0);
#pragma unroll
                for (int f = 0; f < 4; f++)
                    __builtin_amdgcn_global_load_lds((const void *)(src64[f] + (p8 + o64[j])),
                                                     (lds_void *)(buf + f * G::TILE * 8 + el0 * 8), 4, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < G::WF; j++) {
            const int s = wave + j * G::NW;
            if (s < G::SF) {
                __builtin_amdgcn_global_load_lds((const void *)(srcp + (p4 + of[j])),
                                                 (lds_void *)(buf + G::BP + s * 256), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void *)(srcv + (p4 + of[j])),
                                                 (lds_void *)(buf + G::BV + s * 256), 4, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < G::WU; j++) {
            const int s = wave + j * G::NW;
            if (s < G::SU)
                __builtin_amdgcn_global_load_lds((const void *)(srcu + (p4 + ou[j])),
                                                 (lds_void *)(buf + G::BU + s * 256), 4, 0, 0);
        }
    };
    issue(z0, lds);

    // z windows of w, p, mx, my, mz, E (flux type); leading values one plane ahead
    FT ww[9], wp[9], wmx[9], wmy[9], wmz[9], wE[9];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const unsigned o = (unsigned)pmod(z0 - 4 + m, e) * e2 + own;
        ww[m] = q_w[o]; wp[m] = q_p[o];
        wmx[m] = (FT)c_mx[o]; wmy[m] = (FT)c_my[o]; wmz[m] = (FT)c_mz[o]; wE[m] = (FT)c_E[o];
    }
    FT nw, np, lmx, lmy, lmz, lE;
    double nr;  // own rho of the next centre plane
    {
        const unsigned o = (unsigned)pmod(z0 + 4, e) * e2 + own;
        nw = q_w[o]; np = q_p[o];
        lmx = (FT)c_mx[o]; lmy = (FT)c_my[o]; lmz = (FT)c_mz[o]; lE = (FT)c_E[o];
        nr = c_r[(unsigned)z0 * e2 + own];
    }
    unsigned long long cmax = 0ull;
    for (int k = z0; k < z0 + nz; k++) {
        const unsigned zo = (unsigned)k * e2 + own;
        char *const bc = lds + ((k - z0) & 1) * G::BUF;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (k + 1 < z0 + nz) issue(k + 1, lds + (((k - z0) & 1) ^ 1) * G::BUF);
        const double *s_mx = (const double *)bc, *s_my = s_mx + G::TILE, *s_mz = s_my + G::TILE,
                     *s_E = s_mz + G::TILE;
        const FT *s_p = (const FT *)(bc + G::BP), *s_v = (const FT *)(bc + G::BV),
                 *s_u = (const FT *)(bc + G::BU);
        ww[8] = nw; wp[8] = np;
        wmx[8] = lmx; wmy[8] = lmy; wmz[8] = lmz; wE[8] = lE;
        const double rho_c = nr;
        if (k + 1 < z0 + nz) {  // next iteration's operands, hidden behind the compute
            const unsigned o = (unsigned)pmod(k + 5, e) * e2 + own;
            nw = q_w[o]; np = q_p[o];
            lmx = (FT)c_mx[o]; lmy = (FT)c_my[o]; lmz = (FT)c_mz[o]; lE = (FT)c_E[o];
            nr = c_r[zo + e2];
        }
        double rkd[4], rkb[5];
        if (t.active) {  // this cell's RK operands
#pragma unroll
            for (int l = 0; l < 4; l++) rkd[l] = (double)d_c[l][zo];
            if (stage != 1) {
#pragma unroll
                for (int l = 0; l < NC; l++) rkb[l] = b_c[l][zo];
            }
        }
        if (t.active) {
            FT F[5];
#define XMX(m) ((FT)s_mx[me + (m)])
#define XMY(m) ((FT)s_my[me + (m)])
#define XMZ(m) ((FT)s_mz[me + (m)])
#define XE(m) ((FT)s_E[me + (m)])
#define XP(m) s_p[me + (m)]
#define XPF(m) (s_p[me + (m)] + prefF)
#define XU(m) s_u[mu + (m)]
            F[IRHO] = -(KD1(XMX)) * dxi0;
            F[IMX] = -(KHMOMP(XMX, XU, XP)) * dxi0;
            F[IMY] = -(KHMOM(XMY, XU)) * dxi0;
            F[IMZ] = -(KHMOM(XMZ, XU)) * dxi0;
            F[IENE] = -(KHENE(XE, XU, XPF)) * dxi0;
            FENCE();
#define YMX(m) ((FT)s_mx[me + (m) * P])
#define YMY(m) ((FT)s_my[me + (m) * P])
#define YMZ(m) ((FT)s_mz[me + (m) * P])
#define YE(m) ((FT)s_E[me + (m) * P])
#define YP(m) s_p[me + (m) * P]
#define YPF(m) (s_p[me + (m) * P] + prefF)
#define YV(m) s_v[me + (m) * P]
            F[IRHO] = F[IRHO] - (KD1(YMY)) * dxi1;
            F[IMX] = F[IMX] - (KHMOM(YMX, YV)) * dxi1;
            F[IMY] = F[IMY] - (KHMOMP(YMY, YV, YP)) * dxi1;
            F[IMZ] = F[IMZ] - (KHMOM(YMZ, YV)) * dxi1;
            F[IENE] = F[IENE] - (KHENE(YE, YV, YPF)) * dxi1;
            FENCE();
#define ZMX(m) wmx[4 + (m)]
#define ZMY(m) wmy[4 + (m)]
#define ZMZ(m) wmz[4 + (m)]
#define ZE(m) wE[4 + (m)]
#define ZW(m) ww[4 + (m)]
#define ZP(m) wp[4 + (m)]
#define ZPF(m) (wp[4 + (m)] + prefF)
            F[IRHO] = F[IRHO] - (KD1(ZMZ)) * dxi2;
            F[IMX] = F[IMX] - (KHMOM(ZMX, ZW)) * dxi2;
            F[IMY] = F[IMY] - (KHMOM(ZMY, ZW)) * dxi2;
            F[IMZ] = F[IMZ] - (KHMOMP(ZMZ, ZW, ZP)) * dxi2;
            F[IENE] = F[IENE] - (KHENE(ZE, ZW, ZPF)) * dxi2;
            FENCE();
            // stage 1: out = cur + dt*(D+F) (cpu/advance.cpp:82); stages 2 and 3:
            // out = ra*base + rb*(cur + dt*(D+F)) (:114-116, :148-150)
            const double cur_c[5] = {rho_c, s_mx[me], s_my[me], s_mz[me], s_E[me]};
            double o[5];
#pragma unroll
            for (int l = 0; l < NC; l++) {
                const double dl = l == 0 ? 0.0E0 : rkd[l - 1];
                const double df = dl + (double)F[l];
                const double x = cur_c[l] + dt * df;
                o[l] = stage == 1 ? x : ra * rkb[l] + rb * x;
                o_c[l][zo] = o[l];
            }
            double u, v, w, p, T;
            prim_cell(o[0], o[1], o[2], o[3], o[4], u, v, w, p, T);
            qo[zo] = (FT)u;
            qo_v[zo] = (FT)v;
            qo_w[zo] = (FT)w;
            qo_p[zo] = (FT)(p - pref);
            qo_t[zo] = (FT)T;
            if (stage == 3) {
                const unsigned long long cb = cour_bits(o[0], u, v, w, p, dxd);
                cmax = cmax > cb ? cmax : cb;
            }
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            ww[m] = ww[m + 1]; wp[m] = wp[m + 1];
            wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1]; wE[m] = wE[m + 1];
        }
    }
    return cmax;
}

#undef XMX
#undef XMY
#undef XMZ
#undef XE
#undef XP
#undef XPF
#undef XU
#undef YMX
#undef YMY
#undef YMZ
#undef YE
#undef YP
#undef YPF
#undef YV
#undef ZMX
#undef ZMY
#undef ZMZ
#undef ZE
#undef ZW
#undef ZP
#undef ZPF


struct TileArgs {
    int b, x0, y0, z0, nz;
};

__device__ __forceinline__ TileArgs read_tile(const TileDev *__restrict__ tiles, int ti)
{
    const TileDev tl = tiles[ti];
    TileArgs a;
    a.b = __builtin_amdgcn_readfirstlane(tl.box);
    a.x0 = __builtin_amdgcn_readfirstlane(tl.x0);
    a.y0 = __builtin_amdgcn_readfirstlane(tl.y0);
    a.z0 = __builtin_amdgcn_readfirstlane(tl.z0);
    a.nz = __builtin_amdgcn_readfirstlane(tl.nz);
    return a;
}

// One workgroup per tile of one width class of the tile list; the hardware dispatcher
// balances the uneven tiles (partial z runs, stopped boxes) over the CUs -- persistent
// workgroups looping over the list ran 13 % slower on the request mix.  Workgroup b
// takes tile xcd_rank(b), the XCD-contiguous renumbering, so the tiles one XCD works
// on at a time are neighbours in the list (which keeps a box's tiles together) and
// share their halos in that XCD's L2.
__device__ __forceinline__ int xcd_rank(int b, int nb)
{
    const int xcd = b & 7, local = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// The wz field of a stage, over the same tile list as its diffterm launch.
template <typename FT, int LTX>
__global__ void __launch_bounds__(kBatchThreads)
batch_wz_kernel(BoxDev *__restrict__ boxes, const TileDev *__restrict__ tiles, int ntiles, int stage)
{
    const int ti = xcd_rank(blockIdx.x, gridDim.x);
    if (ti >= ntiles) return;
    const TileArgs a = read_tile(tiles, ti);
    if (box_stopped(boxes[a.b])) return;  // uniform over the workgroup
    wz_tile<FT, LTX>(boxes, a.b, a.x0, a.y0, a.z0, a.nz, stage);
}

// Staged diffterm: as many workgroups per CU as its LDS allows (FP32: 32-35 KB for 16-
// and 32-wide tiles, four, with at most 128 VGPRs; 48 KB for 64-wide, three; FP64 twice
// that: two, and one).
template <typename FT, int LTX>
__global__ void __launch_bounds__(kBatchThreads, (160 * 1024) / (DmaGeo<FT, LTX>::LDS_ELEMS * (int)sizeof(FT)) > 4
                                                     ? 4
                                                     : (160 * 1024) / (DmaGeo<FT, LTX>::LDS_ELEMS * (int)sizeof(FT)))
batch_diff_dma_kernel(BoxDev *__restrict__ boxes, const TileDev *__restrict__ tiles, int ntiles, int stage)
{
    __shared__ FT lds[DmaGeo<FT, LTX>::LDS_ELEMS];
    const int ti = xcd_rank(blockIdx.x, gridDim.x);
    if (ti >= ntiles) return;
    const TileArgs a = read_tile(tiles, ti);
    if (box_stopped(boxes[a.b])) return;  // uniform over the workgroup
    diff_tile_dma<FT, LTX>(boxes, a.b, a.x0, a.y0, a.z0, a.nz, stage, lds);
}

template <typename FT, int LTX>
__global__ void __launch_bounds__(kBatchThreads, 2)
batch_diff_kernel(BoxDev *__restrict__ boxes, const TileDev *__restrict__ tiles, int ntiles, int stage)
{
    __shared__ FT lds[9 * kLdsCells];
    const int ti = xcd_rank(blockIdx.x, gridDim.x);
    if (ti >= ntiles) return;
    const TileArgs a = read_tile(tiles, ti);
    if (box_stopped(boxes[a.b])) return;  // uniform over the workgroup
    diff_tile<FT, LTX>(boxes, a.b, a.x0, a.y0, a.z0, a.nz, stage, lds);
}

// The FP32 stage-1 tile (no base operand, no Courant reduction) fits 3 waves per SIMD
// without spilling; the others 2.
template <typename FT, int LTX, int stage, bool KEEP>
__global__ void __launch_bounds__(kBatchThreads, (sizeof(FT) == 4 && stage == 1) ? 3 : 2)
batch_hyp_kernel(BoxDev *__restrict__ boxes, const TileDev *__restrict__ tiles, int ntiles)
{
    __shared__ FT lds[7 * kLdsCells];
    const int ti = xcd_rank(blockIdx.x, gridDim.x);
    if (ti >= ntiles) return;
    const TileArgs a = read_tile(tiles, ti);
    if (box_stopped(boxes[a.b])) return;  // uniform over the workgroup
    const unsigned long long cm = hyp_tile<FT, LTX, stage, KEEP>(boxes, a.b, a.x0, a.y0, a.z0, a.nz, lds);
    if (stage == 3) {
        const unsigned long long m = block_max_u64(cm);
        if (threadIdx.x == 0) atomicMax(&boxes[a.b].acc, m);
    }
}

enum Pass { PASS_WZ, PASS_DIFF, PASS_DIFF_STAGED, PASS_HYP, PASS_HYP_KEEP, PASS_HYP_STAGED };

// 32-wide tiles: the hypterm + update tile staged by direct-to-LDS loads (FP32 flux:
// 53.8 KB of LDS, three workgroups per CU; FP64: 66.6 KB, two).
template <typename FT, int stage>
__global__ void __launch_bounds__(kBatchThreads, 2)
batch_hyp_dma_kernel(BoxDev *__restrict__ boxes, const TileDev *__restrict__ tiles, int ntiles)
{
    __shared__ double lds[2 * HypDmaGeo<FT>::BUF / 8];
    const int ti = xcd_rank(blockIdx.x, gridDim.x);
    if (ti >= ntiles) return;
    const TileArgs a = read_tile(tiles, ti);
    if (box_stopped(boxes[a.b])) return;  // uniform over the workgroup
    const unsigned long long cm = hyp_tile_dma<FT, stage>(boxes, a.b, a.x0, a.y0, a.z0, a.nz, (char *)lds);
    if (stage == 3) {
        const unsigned long long m = block_max_u64(cm);
        if (threadIdx.x == 0) atomicMax(&boxes[a.b].acc, m);
    }
}

template <typename FT, int LTX>
hipError_t stage_class(BoxDev *boxes, const TileDev *tiles, int n, int stage, Pass pass, hipStream_t s)
{
    switch (pass) {
    case PASS_WZ: batch_wz_kernel<FT, LTX><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n, stage); break;
    case PASS_DIFF: batch_diff_kernel<FT, LTX><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n, stage); break;
    case PASS_DIFF_STAGED:
        batch_diff_dma_kernel<FT, LTX><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n, stage);
        break;
    case PASS_HYP_STAGED:  // 32-wide tiles only
        if (LTX != 5) return hipErrorInvalidValue;
        if (stage == 1) batch_hyp_dma_kernel<FT, 1><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n);
        else if (stage == 2) batch_hyp_dma_kernel<FT, 2><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n);
        else batch_hyp_dma_kernel<FT, 3><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n);
        break;
    case PASS_HYP_KEEP:  // stages 2 and 3 of the FP32 flux mode
        if (stage == 2) batch_hyp_kernel<float, LTX, 2, true><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n);
        else batch_hyp_kernel<float, LTX, 3, true><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n);
        break;
    default:
        switch (stage) {
        case 1: batch_hyp_kernel<FT, LTX, 1, false><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n); break;
        case 2: batch_hyp_kernel<FT, LTX, 2, false><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n); break;
        default: batch_hyp_kernel<FT, LTX, 3, false><<<n, kBatchThreads, 0, s>>>(boxes, tiles, n); break;
        }
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_init(BoxDev *boxes, int box, int e, bool fp32, hipStream_t s)
{
    const int64_t e3 = (int64_t)e * e * e;
    int grid = (int)((e3 + kBatchThreads * 4 - 1) / (kBatchThreads * 4));
    grid = grid < 1 ? 1 : (grid > 4096 ? 4096 : grid);
    if (fp32)
        batch_init_kernel<float><<<grid, kBatchThreads, 0, s>>>(boxes, box);
    else
        batch_init_kernel<double><<<grid, kBatchThreads, 0, s>>>(boxes, box);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(BoxDev *boxes, const int *ids, int n, int count_step, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    batch_finalize_kernel<<<(n + 255) / 256, 256, 0, s>>>(boxes, ids, n, count_step);
    return hipGetLastError();
}

hipError_t launch_batch_stage(BoxDev *boxes, const TileDev *tiles, const int start[3],
                              const int count[3], int stage, bool fp32, bool staged_diff,
                              bool keep_centre, bool staged_hyp, bool staged_hyp64, hipStream_t s)
{
    // every class's wz field (staged diffterm only), then every diffterm, then every
    // hypterm + update: each pass reads what the one before wrote on other tiles' cells
    const bool keep = keep_centre && fp32 && stage != 1;
    const Pass passes[3] = {PASS_WZ, staged_diff ? PASS_DIFF_STAGED : PASS_DIFF,
                            keep ? PASS_HYP_KEEP : PASS_HYP};
    for (const Pass pass : passes) {
        if (pass == PASS_WZ && !staged_diff) continue;
        for (int c = 0; c < 3; c++) {
            const int n = count[c];
            if (n <= 0) continue;
            const TileDev *t = tiles + start[c];
            hipError_t e;
            const bool staged = fp32 ? staged_hyp : staged_hyp64;
            const Pass pc = (pass == PASS_HYP && staged && c == 1) ? PASS_HYP_STAGED : pass;
            if (fp32)
                e = c == 0 ? stage_class<float, 4>(boxes, t, n, stage, pc, s)
                  : c == 1 ? stage_class<float, 5>(boxes, t, n, stage, pc, s)
                           : stage_class<float, 6>(boxes, t, n, stage, pc, s);
            else
                e = c == 0 ? stage_class<double, 4>(boxes, t, n, stage, pc, s)
                  : c == 1 ? stage_class<double, 5>(boxes, t, n, stage, pc, s)
                           : stage_class<double, 6>(boxes, t, n, stage, pc, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace cns
