// cns_stencil.hip -- z-marching, LDS-tiled hypterm and diffterm for gfx950.
//
// Tiling (both kernels).  A workgroup owns a 64 x BY column tile of the interior
// (x unit-stride: one 64-lane wave per row, so every global access of a row is a
// coalesced 512-byte burst) and marches up a run of KZ planes.  Each plane's x/y
// neighbourhood (4-cell halos, and for diffterm the edge halos of the mixed
// derivatives) is staged in LDS.  The z neighbours of a lane's own column come
// either from a 9-plane register window (ZWIN bit set) or straight from global
// memory, where the planes the workgroup touched in the last few iterations are
// still in L1/L2; the choice trades registers (occupancy) against L1/L2 traffic.
// Nothing is re-read from HBM except the thin x/y halos, which neighbouring
// workgroups fetch at the same time.  The workgroup -> tile map gives each XCD
// (workgroups are dealt to the 8 XCDs round-robin) a contiguous run of tiles, so
// halos are shared within one L2.
//
// Arithmetic follows the reference's per-cell operation order exactly
// (cpu/hypterm.cpp:11-209, cpu/diffterm.cpp:22-320, cpu/advance.cpp:78-153), so
// with -ffp-contract=off the results are bit-identical to the reference cpu/ path.
#include <cstdlib>

#include "cns_kernels.h"

namespace cns {

namespace {

constexpr int TX = 64;      // tile width (one wave)
constexpr int PX = TX + 8;  // LDS pitch of arrays with x halos

// Workgroup -> tile id such that each XCD (workgroups b with equal b % 8) gets a
// contiguous run of tile ids, i.e. spatially adjacent tiles.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int xcd = b & 7, local = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// The 8*BY x-halo cells (rows 4..BY+3, cols 0..3 and 68..71) and the 8*64 y-halo
// cells (rows 0..3 and BY+4..BY+7, cols 4..67) of an LDS plane, and the 64 corners.
template <int BY>
__device__ __forceinline__ void xhalo_rc(int h, int &r, int &c)
{
    r = 4 + (h >> 3);
    const int cc = h & 7;
    c = cc < 4 ? cc : TX + cc;
}
template <int BY>
__device__ __forceinline__ void yhalo_rc(int h, int &r, int &c)
{
    const int rr = h >> 6;
    c = 4 + (h & 63);
    r = rr < 4 ? rr : BY + rr;
}
template <int BY>
__device__ __forceinline__ void corner_rc(int h, int &r, int &c)
{
    const int rr = h >> 3, cc = h & 7;
    r = rr < 4 ? rr : BY + rr;
    c = cc < 4 ? cc : TX + cc;
}

// Keeps the scheduler from hoisting the loads of later groups of a long
// straight-line block above earlier ones (bounds the live registers).
#define SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)

// First and second 8th-order differences of an offset accessor F(m), before the
// 1/dx factor (cpu/hypterm.cpp:43-47, cpu/diffterm.cpp:77-81 and :159-163).
#define D1(F) (ALP * (F(1) - F(-1)) + BET * (F(2) - F(-2)) + GAM * (F(3) - F(-3)) + DEL * (F(4) - F(-4)))
#define D2(F)                                                                                \
    (CENTER * F(0) + OFF1 * (F(1) + F(-1)) + OFF2 * (F(2) + F(-2)) + OFF3 * (F(3) + F(-3)) \
     + OFF4 * (F(4) + F(-4)))

// Hypterm sums of one direction (cpu/hypterm.cpp x :49-90, y :117-147, z :174-204) for
// momentum flux with the normal velocity UN(m), pressure P(m) and field C(m).
#define HMOM(C, UN) \
    (ALP * (C(1) * UN(1) - C(-1) * UN(-1)) + BET * (C(2) * UN(2) - C(-2) * UN(-2)) \
     + GAM * (C(3) * UN(3) - C(-3) * UN(-3)) + DEL * (C(4) * UN(4) - C(-4) * UN(-4)))
#define HMOMP(C, UN, P)                                                                  \
    (ALP * (C(1) * UN(1) - C(-1) * UN(-1) + (P(1) - P(-1)))                              \
     + BET * (C(2) * UN(2) - C(-2) * UN(-2) + (P(2) - P(-2)))                            \
     + GAM * (C(3) * UN(3) - C(-3) * UN(-3) + (P(3) - P(-3)))                            \
     + DEL * (C(4) * UN(4) - C(-4) * UN(-4) + (P(4) - P(-4))))
#define HENE(C, UN, P)                                                                   \
    (ALP * (C(1) * UN(1) - C(-1) * UN(-1) + (P(1) * UN(1) - P(-1) * UN(-1)))             \
     + BET * (C(2) * UN(2) - C(-2) * UN(-2) + (P(2) * UN(2) - P(-2) * UN(-2)))           \
     + GAM * (C(3) * UN(3) - C(-3) * UN(-3) + (P(3) * UN(3) - P(-3) * UN(-3)))           \
     + DEL * (C(4) * UN(4) - C(-4) * UN(-4) + (P(4) * UN(4) - P(-4) * UN(-4))))

// ======================================================================= hypterm
// ZWIN bits: 1 mx, 2 my, 4 mz, 8 E, 16 w, 32 p -- fields whose z window is in registers.
// With RK the kernel also applies the stage update of cpu/advance.cpp (phase 1..3)
// for its cell, reading D and the stage base and writing `out` (never `cons`).
struct RkArgs {
    const double *d;     // difflux (interior)
    const double *base;  // U^n (ghosted), phases 2 and 3
    double *out;         // result (ghosted); must not alias cons
    const double *dt;    // device dt
    int phase;
};

template <int BY, int WPE, bool RK>
__global__ void __launch_bounds__(TX *BY, WPE)
hypterm_kernel(Geom g, const double *__restrict__ cons, const double *__restrict__ q,
               double *__restrict__ flux, double dxi0, double dxi1, double dxi2, int kz, int ntx,
               int nty, RkArgs rk)
{
    constexpr int NT = TX * BY, R = BY + 8, S = R * PX;
    __shared__ double sm[7 * S];
    double *s_mx = sm, *s_my = sm + S, *s_mz = sm + 2 * S, *s_E = sm + 3 * S, *s_p = sm + 4 * S,
           *s_u = sm + 5 * S, *s_v = sm + 6 * S;

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int gi0 = bx * TX, gj0 = by * BY;  // ghost-shifted coords of LDS cell (0,0)
    const int gi = gi0 + tx + 4, gj = gj0 + ty + 4;
    const bool inbox = gi < g.g0 && gj < g.g1;
    const bool active = gi < g.n0 + NG && gj < g.n1 + NG;
    const int kbeg = bz * kz, kend = min(kbeg + kz, g.n2);
    const int64_t cg = g.cg, pg = g.pg;
    const double *c_mx = cons + IMX * cg, *c_my = cons + IMY * cg, *c_mz = cons + IMZ * cg,
                 *c_E = cons + IENE * cg;
    const double *q_u = q + QU * cg, *q_v = q + QV * cg, *q_w = q + QW * cg, *q_p = q + QPRES * cg;
    const int64_t own = (int64_t)gj * g.xg + gi;
    const int me = (ty + 4) * PX + tx + 4;
    double dt = 0.0, ra = 0.0, rb = 1.0;
    if (RK) {
        // one branch-free form for the three stages: r = ra*base + rb*(cons + dt*(D+F));
        // (ra, rb) = (0, 1), (3/4, 1/4), (1/3, 2/3); 0*base + 1*x is exactly x, so
        // every phase keeps the reference's rounding (cpu/advance.cpp:82, :114-116, :148-150)
        dt = rk.dt[0];
        ra = rk.phase == 1 ? 0.0E0 : (rk.phase == 2 ? ThreeQuarters : OneThird);
        rb = rk.phase == 1 ? 1.0E0 : (rk.phase == 2 ? OneQuarter : TwoThirds);
    }

    // register windows (plane k+m at [4+m]) and the prefetched next leading plane
    double wmx[9], wmy[9], wmz[9], wE[9], ww[9], wp[9];
    double nmx = 0, nmy = 0, nmz = 0, nE = 0, nw = 0, np = 0, nu = 0, nv = 0;
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wmx[m] = c_mx[o]; wmy[m] = c_my[o]; wmz[m] = c_mz[o];
            wE[m] = c_E[o]; ww[m] = q_w[o]; wp[m] = q_p[o];
        }
        const int64_t o = (int64_t)(kbeg + 2 * NG) * pg + own;
        nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o]; nw = q_w[o]; np = q_p[o];
        const int64_t oc = (int64_t)(kbeg + NG) * pg + own;
        nu = q_u[oc];
        nv = q_v[oc];
    }

    for (int k = kbeg; k < kend; k++) {
        const int64_t plane = (int64_t)(k + NG) * pg;
        const int64_t zo = plane + own;  // own cell of the centre plane
        wmx[8] = nmx; wmy[8] = nmy; wmz[8] = nmz; wE[8] = nE; ww[8] = nw; wp[8] = np;
        const double u_own = nu, v_own = nv;
        __syncthreads();
        if (inbox) {
            s_mx[me] = wmx[4]; s_my[me] = wmy[4]; s_mz[me] = wmz[4];
            s_E[me] = wE[4]; s_p[me] = wp[4]; s_u[me] = u_own; s_v[me] = v_own;
        }
        for (int h = threadIdx.x; h < 8 * BY; h += NT) {
            int r, c;
            xhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) {
                const int64_t o = plane + (int64_t)hj * g.xg + hi;
                const int l = r * PX + c;
                s_mx[l] = c_mx[o]; s_my[l] = c_my[o]; s_mz[l] = c_mz[o];
                s_E[l] = c_E[o]; s_p[l] = q_p[o]; s_u[l] = q_u[o];
            }
        }
        for (int h = threadIdx.x; h < 8 * TX; h += NT) {
            int r, c;
            yhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) {
                const int64_t o = plane + (int64_t)hj * g.xg + hi;
                const int l = r * PX + c;
                s_mx[l] = c_mx[o]; s_my[l] = c_my[o]; s_mz[l] = c_mz[o];
                s_E[l] = c_E[o]; s_p[l] = q_p[o]; s_v[l] = q_v[o];
            }
        }
        __syncthreads();
        // Loads whose latency the compute below hides: the leading plane of the next
        // iteration, the next centre u and v, and this cell's RK operands.
        if (inbox && k + 1 < kend) {
            const int64_t o = zo + 5 * pg;
            nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o]; nw = q_w[o]; np = q_p[o];
            nu = q_u[zo + pg];
            nv = q_v[zo + pg];
        }
        const int64_t fo = iidx(g, gi - NG, gj - NG, k);
        double rkd[5], rkb[5], rho_c = 0.0;
        if (RK && active) {
            rho_c = cons[zo];
#pragma unroll
            for (int l = 0; l < NC; l++) {
                rkd[l] = rk.d[l * g.c + fo];
                rkb[l] = rk.base[l * cg + zo];
            }
        }

        if (active) {
            double F[5];
            // ---- x (cpu/hypterm.cpp:33-90): F = -(sum)*dxinv(1)
#define XMX(m) s_mx[me + (m)]
#define XMY(m) s_my[me + (m)]
#define XMZ(m) s_mz[me + (m)]
#define XE(m) s_E[me + (m)]
#define XP(m) s_p[me + (m)]
#define XU(m) s_u[me + (m)]
            F[IRHO] = -(D1(XMX)) * dxi0;
            F[IMX] = -(HMOMP(XMX, XU, XP)) * dxi0;
            F[IMY] = -(HMOM(XMY, XU)) * dxi0;
            F[IMZ] = -(HMOM(XMZ, XU)) * dxi0;
            F[IENE] = -(HENE(XE, XU, XP)) * dxi0;
            // ---- y (cpu/hypterm.cpp:101-147): F = F - (sum)*dxinv(2)
#define YMX(m) s_mx[me + (m) * PX]
#define YMY(m) s_my[me + (m) * PX]
#define YMZ(m) s_mz[me + (m) * PX]
#define YE(m) s_E[me + (m) * PX]
#define YP(m) s_p[me + (m) * PX]
#define YV(m) s_v[me + (m) * PX]
            F[IRHO] = F[IRHO] - (D1(YMY)) * dxi1;
            F[IMX] = F[IMX] - (HMOM(YMX, YV)) * dxi1;
            F[IMY] = F[IMY] - (HMOMP(YMY, YV, YP)) * dxi1;
            F[IMZ] = F[IMZ] - (HMOM(YMZ, YV)) * dxi1;
            F[IENE] = F[IENE] - (HENE(YE, YV, YP)) * dxi1;
            // ---- z (cpu/hypterm.cpp:158-204) from the register windows
#define ZMX(m) wmx[4 + (m)]
#define ZMY(m) wmy[4 + (m)]
#define ZMZ(m) wmz[4 + (m)]
#define ZE(m) wE[4 + (m)]
#define ZW(m) ww[4 + (m)]
#define ZP(m) wp[4 + (m)]
            F[IRHO] = F[IRHO] - (D1(ZMZ)) * dxi2;
            F[IMX] = F[IMX] - (HMOM(ZMX, ZW)) * dxi2;
            F[IMY] = F[IMY] - (HMOM(ZMY, ZW)) * dxi2;
            F[IMZ] = F[IMZ] - (HMOMP(ZMZ, ZW, ZP)) * dxi2;
            F[IENE] = F[IENE] - (HENE(ZE, ZW, ZP)) * dxi2;
#pragma unroll
            for (int l = 0; l < NC; l++) flux[l * g.c + fo] = F[l];
            if (RK) {
                const double uc[5] = {rho_c, wmx[4], wmy[4], wmz[4], wE[4]};
#pragma unroll
                for (int l = 0; l < NC; l++)
                    rk.out[l * cg + zo] = ra * rkb[l] + rb * (uc[l] + dt * (rkd[l] + F[l]));
            }
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1];
            wE[m] = wE[m + 1]; ww[m] = ww[m + 1]; wp[m] = wp[m + 1];
        }
    }
}

// ====================================================================== diffterm
// Per plane k (ghost-shifted kc = k+4) three LDS phases:
//  A  stage the leading plane k+4 (u with x halo, v with y halo) and the centre plane
//     k (u, v with square halos, w and T with cross halos);
//  B  push ux, vy of the leading plane into the register windows; compute the centre
//     plane's ux on the y halo rows, vy on the x halo columns and wz on the cross halo
//     (w of the halo columns at k+-4 is read from global/L2);
//  C  every output of cpu/diffterm.cpp:155-315 for the lane's own cell.
// The reference's nine ghosted temporaries are never materialised; each value it
// would store is computed here by the same expression.
// ZWIN bits: 1 u, 2 v, 4 w, 8 T -- fields whose z window is in registers; ux and vy
// (computed values) always have register windows.
template <int BY>
struct DiffLds {
    static constexpr int R = BY + 8;
    static constexpr int LU = 0;                    // BY x 72   lead u (x halo)
    static constexpr int LV = LU + BY * PX;         // R x 64    lead v (y halo)
    static constexpr int CU = LV + R * TX;          // R x 72    centre u (square)
    static constexpr int CV = CU + R * PX;          // R x 72    centre v (square)
    static constexpr int CW = CV + R * PX;          // R x 72    centre w (cross)
    static constexpr int CT = CW + R * PX;          // R x 72    centre T (cross)
    static constexpr int DUX = CT + R * PX;         // R x 64    ux (tile + y halo rows)
    static constexpr int DVY = DUX + R * TX;        // BY x 72   vy (tile + x halo cols)
    static constexpr int DWZ = DVY + BY * PX;       // R x 72    wz (cross)
    static constexpr int TOTAL = DWZ + R * PX;
};

template <int BY, int WPE, int ZWIN>
__global__ void __launch_bounds__(TX *BY, WPE)
diffterm_kernel(Geom g, const double *__restrict__ q, double *__restrict__ difflux, double dxi0,
                double dxi1, double dxi2, double eta, double alam, int kz, int ntx, int nty)
{
    using L = DiffLds<BY>;
    constexpr int NT = TX * BY;
    __shared__ double sm[L::TOTAL];
    double *s_lu = sm + L::LU, *s_lv = sm + L::LV, *s_cu = sm + L::CU, *s_cv = sm + L::CV,
           *s_cw = sm + L::CW, *s_ct = sm + L::CT, *s_ux = sm + L::DUX, *s_vy = sm + L::DVY,
           *s_wz = sm + L::DWZ;

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int gi0 = bx * TX, gj0 = by * BY;
    const int gi = gi0 + tx + 4, gj = gj0 + ty + 4;
    const bool inbox = gi < g.g0 && gj < g.g1;
    const bool active = gi < g.n0 + NG && gj < g.n1 + NG;
    const int kbeg = bz * kz, kend = min(kbeg + kz, g.n2);
    const int64_t cg = g.cg, pg = g.pg, xg = g.xg;
    const double *q_u = q + QU * cg, *q_v = q + QV * cg, *q_w = q + QW * cg, *q_t = q + QTEMP * cg;
    const int64_t own = (int64_t)gj * xg + gi;
    const int me = (ty + 4) * PX + tx + 4;     // own cell in R x 72 arrays
    const int me64 = (ty + 4) * TX + tx;       // own cell in R x 64 arrays
    const int meby = ty * PX + tx + 4;         // own cell in BY x 72 arrays
    const double dxi0s = dxi0 * dxi0, dxi1s = dxi1 * dxi1, dxi2s = dxi2 * dxi2;

    // windows: plane k+m at [4+m]
    double wu[9], wv[9], ww[9], wt[9], wux[9], wvy[9];
    if (ZWIN && inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            if (ZWIN & 1) wu[m] = q_u[o];
            if (ZWIN & 2) wv[m] = q_v[o];
            if (ZWIN & 4) ww[m] = q_w[o];
            if (ZWIN & 8) wt[m] = q_t[o];
        }
    }
#define LUX(mm) s_lu[meby + (mm)]
#define LVY(mm) s_lv[me64 + (mm) * TX]
    // ux, vy of the first 8 planes of the window: staged through LDS one plane at a
    // time (a rolled loop: unrolled, its barriers make the compiler keep every
    // plane's loads live at once)
#pragma unroll 1
    for (int m = 0; m < 8; m++) {
        const int64_t plane = (int64_t)(kbeg + m) * pg;
        __syncthreads();
        if (inbox) {
            s_lu[meby] = q_u[plane + own];
            s_lv[me64] = q_v[plane + own];
        }
        for (int h = threadIdx.x; h < 8 * BY; h += NT) {
            int r, c;
            xhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) s_lu[(r - 4) * PX + c] = q_u[plane + (int64_t)hj * xg + hi];
        }
        for (int h = threadIdx.x; h < 8 * TX; h += NT) {
            int r, c;
            yhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) s_lv[r * TX + c - 4] = q_v[plane + (int64_t)hj * xg + hi];
        }
        __syncthreads();
        {
            // shift in from the top so that after 8 planes slot m holds plane kbeg+m
            const double ux_ = D1(LUX) * dxi0, vy_ = D1(LVY) * dxi1;
#pragma unroll
            for (int mm = 0; mm < 7; mm++) {
                wux[mm] = wux[mm + 1];
                wvy[mm] = wvy[mm + 1];
            }
            wux[7] = ux_;
            wvy[7] = vy_;
        }
    }

    // prefetched own values of the next leading plane
    double nu = 0.0, nv = 0.0, nw = 0.0, nt = 0.0;
    if (inbox) {
        const int64_t o = (int64_t)(kbeg + 2 * NG) * pg + own;
        nu = q_u[o]; nv = q_v[o]; nw = q_w[o]; nt = q_t[o];
    }
    for (int k = kbeg; k < kend; k++) {
        const int64_t cplane = (int64_t)(k + NG) * pg;
        const int64_t lplane = (int64_t)(k + 2 * NG) * pg;
        const int64_t zo = cplane + own;
        const double lead_u = nu, lead_v = nv;
        double cen[4] = {0, 0, 0, 0};
        if (inbox) {
            if (ZWIN & 1) wu[8] = nu;
            if (ZWIN & 2) wv[8] = nv;
            if (ZWIN & 4) ww[8] = nw;
            if (ZWIN & 8) wt[8] = nt;
            cen[0] = (ZWIN & 1) ? wu[4] : q_u[zo];
            cen[1] = (ZWIN & 2) ? wv[4] : q_v[zo];
            cen[2] = (ZWIN & 4) ? ww[4] : q_w[zo];
            cen[3] = (ZWIN & 8) ? wt[4] : q_t[zo];
        }
        // ---------------- phase A
        __syncthreads();
        if (inbox) {
            s_lu[meby] = lead_u;
            s_lv[me64] = lead_v;
            s_cu[me] = cen[0]; s_cv[me] = cen[1]; s_cw[me] = cen[2]; s_ct[me] = cen[3];
        }
        for (int h = threadIdx.x; h < 8 * BY; h += NT) {
            int r, c;
            xhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) {
                const int64_t ho = (int64_t)hj * xg + hi;
                const int l = r * PX + c;
                s_lu[(r - 4) * PX + c] = q_u[lplane + ho];
                s_cu[l] = q_u[cplane + ho]; s_cv[l] = q_v[cplane + ho];
                s_cw[l] = q_w[cplane + ho]; s_ct[l] = q_t[cplane + ho];
            }
        }
        for (int h = threadIdx.x; h < 8 * TX; h += NT) {
            int r, c;
            yhalo_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) {
                const int64_t ho = (int64_t)hj * xg + hi;
                const int l = r * PX + c;
                s_lv[r * TX + c - 4] = q_v[lplane + ho];
                s_cu[l] = q_u[cplane + ho]; s_cv[l] = q_v[cplane + ho];
                s_cw[l] = q_w[cplane + ho]; s_ct[l] = q_t[cplane + ho];
            }
        }
        for (int h = threadIdx.x; h < 64; h += NT) {
            int r, c;
            corner_rc<BY>(h, r, c);
            const int hi = gi0 + c, hj = gj0 + r;
            if (hi < g.g0 && hj < g.g1) {
                const int64_t ho = cplane + (int64_t)hj * xg + hi;
                const int l = r * PX + c;
                s_cu[l] = q_u[ho]; s_cv[l] = q_v[ho];
            }
        }
        __syncthreads();

        // ---------------- phase B
        {
            // leading plane: ux, vy of the own column (cpu/diffterm.cpp:77-81, :110-114)
            wux[8] = D1(LUX) * dxi0;
            wvy[8] = D1(LVY) * dxi1;
            s_ux[me64] = wux[4];
            s_vy[meby] = wvy[4];
            // own wz (cpu/diffterm.cpp:143-147)
#define ZWW(mm) ((ZWIN & 4) ? ww[4 + (mm)] : q_w[zo + (int64_t)(mm) * pg])
            s_wz[me] = D1(ZWW) * dxi2;
            // ux on the y-halo rows (x derivative of the square u tile)
            for (int h = threadIdx.x; h < 8 * TX; h += NT) {
                int r, c;
                yhalo_rc<BY>(h, r, c);
                const int l = r * PX + c;
#define CUX(mm) s_cu[l + (mm)]
                s_ux[r * TX + c - 4] = D1(CUX) * dxi0;
#undef CUX
            }
            // vy on the x-halo columns (y derivative of the square v tile)
            for (int h = threadIdx.x; h < 8 * BY; h += NT) {
                int r, c;
                xhalo_rc<BY>(h, r, c);
                const int l = r * PX + c;
#define CVY(mm) s_cv[l + (mm) * PX]
                s_vy[(r - 4) * PX + c] = D1(CVY) * dxi1;
#undef CVY
            }
            // wz on the cross halo, from w at k+-4 of the halo columns
            for (int h = threadIdx.x; h < 8 * BY + 8 * TX; h += NT) {
                int r, c;
                if (h < 8 * BY) xhalo_rc<BY>(h, r, c);
                else yhalo_rc<BY>(h - 8 * BY, r, c);
                const int hi = gi0 + c, hj = gj0 + r;
                if (hi < g.g0 && hj < g.g1) {
                    const double *wc = q_w + cplane + (int64_t)hj * xg + hi;
#define GWZ(mm) wc[(int64_t)(mm) * pg]
                    s_wz[r * PX + c] = D1(GWZ) * dxi2;
#undef GWZ
                }
            }
        }
        __syncthreads();
        // the next iteration's leading plane: its latency hides behind phase C
        if (inbox && k + 1 < kend) {
            const int64_t o = lplane + pg + own;
            nu = q_u[o]; nv = q_v[o]; nw = q_w[o]; nt = q_t[o];
        }

        // ---------------- phase C
        if (active) {
#define XS(a, mm) a[me + (mm)]
#define YS(a, mm) a[me + (mm) * PX]
#define CUx(mm) XS(s_cu, mm)
#define CUy(mm) YS(s_cu, mm)
#define CUz(mm) ((ZWIN & 1) ? wu[4 + (mm)] : q_u[zo + (int64_t)(mm) * pg])
#define CVx(mm) XS(s_cv, mm)
#define CVy(mm) YS(s_cv, mm)
#define CVz(mm) ((ZWIN & 2) ? wv[4 + (mm)] : q_v[zo + (int64_t)(mm) * pg])
#define CWx(mm) XS(s_cw, mm)
#define CWy(mm) YS(s_cw, mm)
#define CWz(mm) ZWW(mm)
#define CTx(mm) XS(s_ct, mm)
#define CTy(mm) YS(s_ct, mm)
#define CTz(mm) ((ZWIN & 8) ? wt[4 + (mm)] : q_t[zo + (int64_t)(mm) * pg])
#define VYx(mm) s_vy[meby + (mm)]
#define WZx(mm) XS(s_wz, mm)
#define UXy(mm) s_ux[me64 + (mm) * TX]
#define WZy(mm) YS(s_wz, mm)
#define UXz(mm) wux[4 + (mm)]
#define VYz(mm) wvy[4 + (mm)]
            // cpu/diffterm.cpp:159-187
            const double uxx = D2(CUx) * dxi0s;
            const double uyy = D2(CUy) * dxi1s;
            const double uzz = D2(CUz) * dxi2s;
            const double vyx = D1(VYx) * dxi0;
            const double wzx = D1(WZx) * dxi0;
            const double dmx = eta * (FourThirds * uxx + uyy + uzz + OneThird * (vyx + wzx));
            SCHED_FENCE();
            // :198-226
            const double vxx = D2(CVx) * dxi0s;
            const double vyy = D2(CVy) * dxi1s;
            const double vzz = D2(CVz) * dxi2s;
            const double uxy = D1(UXy) * dxi1;
            const double wzy = D1(WZy) * dxi1;
            const double dmy = eta * (vxx + FourThirds * vyy + vzz + OneThird * (uxy + wzy));
            SCHED_FENCE();
            // :237-265
            const double wxx = D2(CWx) * dxi0s;
            const double wyy = D2(CWy) * dxi1s;
            const double wzz = D2(CWz) * dxi2s;
            const double uxz = D1(UXz) * dxi2;
            const double vyz = D1(VYz) * dxi2;
            const double dmz = eta * (wxx + wyy + FourThirds * wzz + OneThird * (uxz + vyz));
            SCHED_FENCE();
            // :276-311
            const double txx = D2(CTx) * dxi0s;
            const double tyy = D2(CTy) * dxi1s;
            const double tzz = D2(CTz) * dxi2s;
            SCHED_FENCE();
            const double ux = wux[4], vy = wvy[4], wz = s_wz[me];
            const double uy = D1(CUy) * dxi1, uz = D1(CUz) * dxi2;
            const double vx = D1(CVx) * dxi0, vz = D1(CVz) * dxi2;
            const double wx = D1(CWx) * dxi0, wy = D1(CWy) * dxi1;
            const double divu = TwoThirds * (ux + vy + wz);
            const double tauxx = 2.E0 * ux - divu;
            const double tauyy = 2.E0 * vy - divu;
            const double tauzz = 2.E0 * wz - divu;
            const double tauxy = uy + vx;
            const double tauxz = uz + wx;
            const double tauyz = vz + wy;
            double mechwork = tauxx * ux + tauyy * vy + tauzz * wz + (tauxy * tauxy) + (tauxz * tauxz)
                              + (tauyz * tauyz);
            mechwork = eta * mechwork + dmx * cen[0] + dmy * cen[1] + dmz * cen[2];
            const int64_t fo = iidx(g, gi - NG, gj - NG, k);
            const int64_t c = g.c;
            difflux[fo] = 0.0E0;
            difflux[fo + IMX * c] = dmx;
            difflux[fo + IMY * c] = dmy;
            difflux[fo + IMZ * c] = dmz;
            difflux[fo + IENE * c] = alam * (txx + tyy + tzz) + mechwork;
#undef XS
#undef YS
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
            if (ZWIN & 1) wu[m] = wu[m + 1];
            if (ZWIN & 2) wv[m] = wv[m + 1];
            if (ZWIN & 4) ww[m] = ww[m + 1];
            if (ZWIN & 8) wt[m] = wt[m + 1];
            wux[m] = wux[m + 1];
            wvy[m] = wvy[m + 1];
        }
    }
#undef ZWW
#undef LUX
#undef LVY
}

int pick_kz(const Geom &g, int tiles_xy)
{
    // Enough workgroups to fill 256 CUs several times over, with runs of planes
    // long enough that the 8-plane window warm-up stays a small overhead.
    int kz = g.n2;
    while (kz > 16 && (int64_t)tiles_xy * ((g.n2 + kz - 1) / kz) < 2048) kz = (kz + 1) / 2;
    return kz;
}

template <int BY, int WPE, bool RK>
hipError_t run_hyp(const Geom &g, const double *cons, const double *q, double *flux, double dxi0,
                   double dxi1, double dxi2, const RkArgs &rk, hipStream_t s)
{
    const int ntx = (g.n0 + TX - 1) / TX, nty = (g.n1 + BY - 1) / BY;
    const int kz = pick_kz(g, ntx * nty);
    const int ntz = (g.n2 + kz - 1) / kz;
    hypterm_kernel<BY, WPE, RK><<<ntx * nty * ntz, TX * BY, 0, s>>>(
        g, cons, q, flux, dxi0, dxi1, dxi2, kz, ntx, nty, rk);
    return hipGetLastError();
}

template <int BY, int WPE, int ZWIN>
hipError_t run_diff(const Geom &g, const double *q, double *difflux, double dxi0, double dxi1,
                    double dxi2, double eta, double alam, hipStream_t s)
{
    const int ntx = (g.n0 + TX - 1) / TX, nty = (g.n1 + BY - 1) / BY;
    const int kz = pick_kz(g, ntx * nty);
    const int ntz = (g.n2 + kz - 1) / kz;
    diffterm_kernel<BY, WPE, ZWIN><<<ntx * nty * ntz, TX * BY, 0, s>>>(
        g, q, difflux, dxi0, dxi1, dxi2, eta, alam, kz, ntx, nty);
    return hipGetLastError();
}

int env_variant(const char *name)
{
    const char *v = getenv(name);
    return v ? atoi(v) : 0;
}

template <bool RK>
hipError_t hyp_dispatch(const Geom &g, const double *cons, const double *q, double *flux,
                        double dxi0, double dxi1, double dxi2, const RkArgs &rk, hipStream_t s)
{
    switch (env_variant("CNS_HYP_VARIANT")) {
    case 1: return run_hyp<4, 3, RK>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
    case 2: return run_hyp<8, 2, RK>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
    case 3: return run_hyp<4, 1, RK>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
    default: return run_hyp<4, 2, RK>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
    }
}

}  // namespace

hipError_t launch_hypterm_tiled(const Geom &g, const double *cons, const double *q, double *flux,
                                double dxi0, double dxi1, double dxi2, hipStream_t s)
{
    RkArgs rk{nullptr, nullptr, nullptr, nullptr, 0};
    return hyp_dispatch<false>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
}

hipError_t launch_hypterm_rk(const Geom &g, const double *cons, const double *q, double *flux,
                             double dxi0, double dxi1, double dxi2, const double *d,
                             const double *base, double *out, const double *dt_dev, int phase,
                             hipStream_t s)
{
    RkArgs rk{d, base, out, dt_dev, phase};
    return hyp_dispatch<true>(g, cons, q, flux, dxi0, dxi1, dxi2, rk, s);
}

hipError_t launch_diffterm_tiled(const Geom &g, const double *q, double *difflux, double dxi0,
                                 double dxi1, double dxi2, double eta, double alam, hipStream_t s)
{
    switch (env_variant("CNS_DIFF_VARIANT")) {
    case 1: return run_diff<4, 1, 15>(g, q, difflux, dxi0, dxi1, dxi2, eta, alam, s);
    case 2: return run_diff<8, 2, 15>(g, q, difflux, dxi0, dxi1, dxi2, eta, alam, s);
    case 3: return run_diff<4, 3, 15>(g, q, difflux, dxi0, dxi1, dxi2, eta, alam, s);
    default: return run_diff<4, 2, 15>(g, q, difflux, dxi0, dxi1, dxi2, eta, alam, s);
    }
}

}  // namespace cns
