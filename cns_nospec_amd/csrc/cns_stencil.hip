// cns_stencil.hip -- z-marching, LDS-tiled hypterm and diffterm for gfx950.
//
// Tiling (both kernels).  A workgroup owns a 64 x BY column tile of the interior
// (x unit-stride: one 64-lane wave per row, so every global access of a row is a
// coalesced 512-byte burst) and marches up a run of KZ planes.  Each lane keeps a
// 9-plane register window of the fields its own column needs for the z
// derivatives; each plane's x/y neighbourhood (4-cell halos, and for diffterm the
// edge halos of the mixed derivatives) is staged in LDS.  Nothing is re-read
// from HBM except the thin halos, which neighbouring workgroups fetch at the same
// time and so come from L2.  The workgroup -> tile map groups neighbouring tiles
// on one XCD (workgroups are dealt to the 8 XCDs round-robin), so halos are
// shared within one L2.
//
// Arithmetic follows the reference's per-cell operation order exactly
// (cpu/hypterm.cpp:11-209, cpu/diffterm.cpp:22-320), so with -ffp-contract=off
// the results are bit-identical to the reference cpu/ path.
#include "cns_kernels.h"

namespace cns {

namespace {

constexpr int TX = 64;        // tile width (one wave)
constexpr int PX = TX + 8;    // LDS pitch of arrays with x halos

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

// ======================================================================= hypterm
// LDS fields of one plane: mx, my, mz, E, p (x and y halos), u (x halo), v (y halo).
template <int BY>
__global__ void __launch_bounds__(TX *BY)
hypterm_tiled_kernel(Geom g, const double *__restrict__ cons, const double *__restrict__ q,
                     double *__restrict__ flux, double dxi0, double dxi1, double dxi2, int kz,
                     int ntx, int nty)
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

    double wmx[9], wmy[9], wmz[9], wE[9], ww[9], wp[9];
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wmx[m] = c_mx[o]; wmy[m] = c_my[o]; wmz[m] = c_mz[o];
            wE[m] = c_E[o]; ww[m] = q_w[o]; wp[m] = q_p[o];
        }
    }

    for (int k = kbeg; k < kend; k++) {
        const int64_t plane = (int64_t)(k + NG) * pg;
        double u_own = 0.0, v_own = 0.0;
        if (inbox) {
            const int64_t o = (int64_t)(k + 2 * NG) * pg + own;
            wmx[8] = c_mx[o]; wmy[8] = c_my[o]; wmz[8] = c_mz[o];
            wE[8] = c_E[o]; ww[8] = q_w[o]; wp[8] = q_p[o];
            u_own = q_u[plane + own];
            v_own = q_v[plane + own];
        }
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

        if (active) {
            double F[5];
            // ---- x (cpu/hypterm.cpp:33-90): F = -(sum)*dxinv(1)
            {
#define XC(a, m) a[me + (m)]
                const double u1 = XC(s_u, 1), u2 = XC(s_u, 2), u3 = XC(s_u, 3), u4 = XC(s_u, 4);
                const double um1 = XC(s_u, -1), um2 = XC(s_u, -2), um3 = XC(s_u, -3), um4 = XC(s_u, -4);
                F[IRHO] = -(ALP * (XC(s_mx, 1) - XC(s_mx, -1)) + BET * (XC(s_mx, 2) - XC(s_mx, -2))
                            + GAM * (XC(s_mx, 3) - XC(s_mx, -3)) + DEL * (XC(s_mx, 4) - XC(s_mx, -4))) * dxi0;
                F[IMX] = -(ALP * (XC(s_mx, 1) * u1 - XC(s_mx, -1) * um1 + (XC(s_p, 1) - XC(s_p, -1)))
                           + BET * (XC(s_mx, 2) * u2 - XC(s_mx, -2) * um2 + (XC(s_p, 2) - XC(s_p, -2)))
                           + GAM * (XC(s_mx, 3) * u3 - XC(s_mx, -3) * um3 + (XC(s_p, 3) - XC(s_p, -3)))
                           + DEL * (XC(s_mx, 4) * u4 - XC(s_mx, -4) * um4 + (XC(s_p, 4) - XC(s_p, -4)))) * dxi0;
                F[IMY] = -(ALP * (XC(s_my, 1) * u1 - XC(s_my, -1) * um1)
                           + BET * (XC(s_my, 2) * u2 - XC(s_my, -2) * um2)
                           + GAM * (XC(s_my, 3) * u3 - XC(s_my, -3) * um3)
                           + DEL * (XC(s_my, 4) * u4 - XC(s_my, -4) * um4)) * dxi0;
                F[IMZ] = -(ALP * (XC(s_mz, 1) * u1 - XC(s_mz, -1) * um1)
                           + BET * (XC(s_mz, 2) * u2 - XC(s_mz, -2) * um2)
                           + GAM * (XC(s_mz, 3) * u3 - XC(s_mz, -3) * um3)
                           + DEL * (XC(s_mz, 4) * u4 - XC(s_mz, -4) * um4)) * dxi0;
                F[IENE] = -(ALP * (XC(s_E, 1) * u1 - XC(s_E, -1) * um1 + (XC(s_p, 1) * u1 - XC(s_p, -1) * um1))
                            + BET * (XC(s_E, 2) * u2 - XC(s_E, -2) * um2 + (XC(s_p, 2) * u2 - XC(s_p, -2) * um2))
                            + GAM * (XC(s_E, 3) * u3 - XC(s_E, -3) * um3 + (XC(s_p, 3) * u3 - XC(s_p, -3) * um3))
                            + DEL * (XC(s_E, 4) * u4 - XC(s_E, -4) * um4 + (XC(s_p, 4) * u4 - XC(s_p, -4) * um4))) * dxi0;
#undef XC
            }
            // ---- y (cpu/hypterm.cpp:101-147): F = F - (sum)*dxinv(2)
            {
#define YC(a, m) a[me + (m) * PX]
                const double v1 = YC(s_v, 1), v2 = YC(s_v, 2), v3 = YC(s_v, 3), v4 = YC(s_v, 4);
                const double vm1 = YC(s_v, -1), vm2 = YC(s_v, -2), vm3 = YC(s_v, -3), vm4 = YC(s_v, -4);
                F[IRHO] = F[IRHO] - (ALP * (YC(s_my, 1) - YC(s_my, -1)) + BET * (YC(s_my, 2) - YC(s_my, -2))
                                     + GAM * (YC(s_my, 3) - YC(s_my, -3)) + DEL * (YC(s_my, 4) - YC(s_my, -4))) * dxi1;
                F[IMX] = F[IMX] - (ALP * (YC(s_mx, 1) * v1 - YC(s_mx, -1) * vm1)
                                   + BET * (YC(s_mx, 2) * v2 - YC(s_mx, -2) * vm2)
                                   + GAM * (YC(s_mx, 3) * v3 - YC(s_mx, -3) * vm3)
                                   + DEL * (YC(s_mx, 4) * v4 - YC(s_mx, -4) * vm4)) * dxi1;
                F[IMY] = F[IMY] - (ALP * (YC(s_my, 1) * v1 - YC(s_my, -1) * vm1 + (YC(s_p, 1) - YC(s_p, -1)))
                                   + BET * (YC(s_my, 2) * v2 - YC(s_my, -2) * vm2 + (YC(s_p, 2) - YC(s_p, -2)))
                                   + GAM * (YC(s_my, 3) * v3 - YC(s_my, -3) * vm3 + (YC(s_p, 3) - YC(s_p, -3)))
                                   + DEL * (YC(s_my, 4) * v4 - YC(s_my, -4) * vm4 + (YC(s_p, 4) - YC(s_p, -4)))) * dxi1;
                F[IMZ] = F[IMZ] - (ALP * (YC(s_mz, 1) * v1 - YC(s_mz, -1) * vm1)
                                   + BET * (YC(s_mz, 2) * v2 - YC(s_mz, -2) * vm2)
                                   + GAM * (YC(s_mz, 3) * v3 - YC(s_mz, -3) * vm3)
                                   + DEL * (YC(s_mz, 4) * v4 - YC(s_mz, -4) * vm4)) * dxi1;
                F[IENE] = F[IENE] - (ALP * (YC(s_E, 1) * v1 - YC(s_E, -1) * vm1 + (YC(s_p, 1) * v1 - YC(s_p, -1) * vm1))
                                     + BET * (YC(s_E, 2) * v2 - YC(s_E, -2) * vm2 + (YC(s_p, 2) * v2 - YC(s_p, -2) * vm2))
                                     + GAM * (YC(s_E, 3) * v3 - YC(s_E, -3) * vm3 + (YC(s_p, 3) * v3 - YC(s_p, -3) * vm3))
                                     + DEL * (YC(s_E, 4) * v4 - YC(s_E, -4) * vm4 + (YC(s_p, 4) * v4 - YC(s_p, -4) * vm4))) * dxi1;
#undef YC
            }
            // ---- z (cpu/hypterm.cpp:158-204) from the register window, plane k+m at [4+m]
            {
#define ZC(a, m) a[4 + (m)]
                F[IRHO] = F[IRHO] - (ALP * (ZC(wmz, 1) - ZC(wmz, -1)) + BET * (ZC(wmz, 2) - ZC(wmz, -2))
                                     + GAM * (ZC(wmz, 3) - ZC(wmz, -3)) + DEL * (ZC(wmz, 4) - ZC(wmz, -4))) * dxi2;
                F[IMX] = F[IMX] - (ALP * (ZC(wmx, 1) * ZC(ww, 1) - ZC(wmx, -1) * ZC(ww, -1))
                                   + BET * (ZC(wmx, 2) * ZC(ww, 2) - ZC(wmx, -2) * ZC(ww, -2))
                                   + GAM * (ZC(wmx, 3) * ZC(ww, 3) - ZC(wmx, -3) * ZC(ww, -3))
                                   + DEL * (ZC(wmx, 4) * ZC(ww, 4) - ZC(wmx, -4) * ZC(ww, -4))) * dxi2;
                F[IMY] = F[IMY] - (ALP * (ZC(wmy, 1) * ZC(ww, 1) - ZC(wmy, -1) * ZC(ww, -1))
                                   + BET * (ZC(wmy, 2) * ZC(ww, 2) - ZC(wmy, -2) * ZC(ww, -2))
                                   + GAM * (ZC(wmy, 3) * ZC(ww, 3) - ZC(wmy, -3) * ZC(ww, -3))
                                   + DEL * (ZC(wmy, 4) * ZC(ww, 4) - ZC(wmy, -4) * ZC(ww, -4))) * dxi2;
                F[IMZ] = F[IMZ] - (ALP * (ZC(wmz, 1) * ZC(ww, 1) - ZC(wmz, -1) * ZC(ww, -1) + (ZC(wp, 1) - ZC(wp, -1)))
                                   + BET * (ZC(wmz, 2) * ZC(ww, 2) - ZC(wmz, -2) * ZC(ww, -2) + (ZC(wp, 2) - ZC(wp, -2)))
                                   + GAM * (ZC(wmz, 3) * ZC(ww, 3) - ZC(wmz, -3) * ZC(ww, -3) + (ZC(wp, 3) - ZC(wp, -3)))
                                   + DEL * (ZC(wmz, 4) * ZC(ww, 4) - ZC(wmz, -4) * ZC(ww, -4) + (ZC(wp, 4) - ZC(wp, -4)))) * dxi2;
                F[IENE] = F[IENE] - (ALP * (ZC(wE, 1) * ZC(ww, 1) - ZC(wE, -1) * ZC(ww, -1) + (ZC(wp, 1) * ZC(ww, 1) - ZC(wp, -1) * ZC(ww, -1)))
                                     + BET * (ZC(wE, 2) * ZC(ww, 2) - ZC(wE, -2) * ZC(ww, -2) + (ZC(wp, 2) * ZC(ww, 2) - ZC(wp, -2) * ZC(ww, -2)))
                                     + GAM * (ZC(wE, 3) * ZC(ww, 3) - ZC(wE, -3) * ZC(ww, -3) + (ZC(wp, 3) * ZC(ww, 3) - ZC(wp, -3) * ZC(ww, -3)))
                                     + DEL * (ZC(wE, 4) * ZC(ww, 4) - ZC(wE, -4) * ZC(ww, -4) + (ZC(wp, 4) * ZC(ww, 4) - ZC(wp, -4) * ZC(ww, -4)))) * dxi2;
#undef ZC
            }
            const int64_t fo = iidx(g, gi - NG, gj - NG, k);
#pragma unroll
            for (int l = 0; l < NC; l++) flux[l * g.c + fo] = F[l];
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1];
            wE[m] = wE[m + 1]; ww[m] = ww[m + 1]; wp[m] = wp[m + 1];
        }
    }
}

// ====================================================================== diffterm
// First and second 8th-order differences of a value accessor.
#define D1(F) (ALP * (F(1) - F(-1)) + BET * (F(2) - F(-2)) + GAM * (F(3) - F(-3)) + DEL * (F(4) - F(-4)))
#define D2(F)                                                                                \
    (CENTER * F(0) + OFF1 * (F(1) + F(-1)) + OFF2 * (F(2) + F(-2)) + OFF3 * (F(3) + F(-3)) \
     + OFF4 * (F(4) + F(-4)))

// Per plane k (ghost-shifted kc = k+4) three LDS phases:
//  A  stage the leading plane k+4 (u with x halo, v with y halo) and the centre plane
//     k (u, v with square halos, w and T with cross halos);
//  B  push ux, vy of the leading plane into the register windows; compute the centre
//     plane's ux on the y halo rows, vy on the x halo columns and wz on the cross halo
//     (w of the halo columns at k+-4 is read from global/L2);
//  C  every output of cpu/diffterm.cpp:155-315 for the lane's own cell.
// The reference's nine ghosted temporaries are never materialised; each value it
// would store is computed here by the same expression.
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

template <int BY>
__global__ void __launch_bounds__(TX *BY)
diffterm_tiled_kernel(Geom g, const double *__restrict__ q, double *__restrict__ difflux, double dxi0,
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
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wu[m] = q_u[o]; wv[m] = q_v[o]; ww[m] = q_w[o]; wt[m] = q_t[o];
        }
    }
    // ux, vy of the first 8 planes of the window: staged through LDS one plane at a time
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
#define LUX(mm) s_lu[meby + (mm)]
#define LVY(mm) s_lv[me64 + (mm) * TX]
        const double uxm = D1(LUX) * dxi0;
        const double vym = D1(LVY) * dxi1;
#pragma unroll
        for (int mm = 0; mm < 8; mm++)
            if (mm == m) { wux[mm] = uxm; wvy[mm] = vym; }
    }

    for (int k = kbeg; k < kend; k++) {
        const int64_t cplane = (int64_t)(k + NG) * pg;
        const int64_t lplane = (int64_t)(k + 2 * NG) * pg;
        if (inbox) {
            const int64_t o = lplane + own;
            wu[8] = q_u[o]; wv[8] = q_v[o]; ww[8] = q_w[o]; wt[8] = q_t[o];
        }
        // ---------------- phase A
        __syncthreads();
        if (inbox) {
            s_lu[meby] = wu[8];
            s_lv[me64] = wv[8];
            s_cu[me] = wu[4]; s_cv[me] = wv[4]; s_cw[me] = ww[4]; s_ct[me] = wt[4];
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
#undef LUX
#undef LVY
            // centre plane own values
            s_ux[me64] = wux[4];
            s_vy[meby] = wvy[4];
#define WW(mm) ww[4 + (mm)]
            s_wz[me] = D1(WW) * dxi2;  // cpu/diffterm.cpp:143-147
#undef WW
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

        // ---------------- phase C
        if (active) {
#define XS(a, mm) a[me + (mm)]
#define YS(a, mm) a[me + (mm) * PX]
#define ZW(a, mm) a[4 + (mm)]
#define CUx(mm) XS(s_cu, mm)
#define CUy(mm) YS(s_cu, mm)
#define CUz(mm) ZW(wu, mm)
#define CVx(mm) XS(s_cv, mm)
#define CVy(mm) YS(s_cv, mm)
#define CVz(mm) ZW(wv, mm)
#define CWx(mm) XS(s_cw, mm)
#define CWy(mm) YS(s_cw, mm)
#define CWz(mm) ZW(ww, mm)
#define CTx(mm) XS(s_ct, mm)
#define CTy(mm) YS(s_ct, mm)
#define CTz(mm) ZW(wt, mm)
#define VYx(mm) s_vy[meby + (mm)]
#define WZx(mm) XS(s_wz, mm)
#define UXy(mm) s_ux[me64 + (mm) * TX]
#define WZy(mm) YS(s_wz, mm)
#define UXz(mm) ZW(wux, mm)
#define VYz(mm) ZW(wvy, mm)
            // cpu/diffterm.cpp:159-187
            const double uxx = D2(CUx) * dxi0s;
            const double uyy = D2(CUy) * dxi1s;
            const double uzz = D2(CUz) * dxi2s;
            const double vyx = D1(VYx) * dxi0;
            const double wzx = D1(WZx) * dxi0;
            const double dmx = eta * (FourThirds * uxx + uyy + uzz + OneThird * (vyx + wzx));
            // :198-226
            const double vxx = D2(CVx) * dxi0s;
            const double vyy = D2(CVy) * dxi1s;
            const double vzz = D2(CVz) * dxi2s;
            const double uxy = D1(UXy) * dxi1;
            const double wzy = D1(WZy) * dxi1;
            const double dmy = eta * (vxx + FourThirds * vyy + vzz + OneThird * (uxy + wzy));
            // :237-265
            const double wxx = D2(CWx) * dxi0s;
            const double wyy = D2(CWy) * dxi1s;
            const double wzz = D2(CWz) * dxi2s;
            const double uxz = D1(UXz) * dxi2;
            const double vyz = D1(VYz) * dxi2;
            const double dmz = eta * (wxx + wyy + FourThirds * wzz + OneThird * (uxz + vyz));
            // :276-311
            const double txx = D2(CTx) * dxi0s;
            const double tyy = D2(CTy) * dxi1s;
            const double tzz = D2(CTz) * dxi2s;
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
            mechwork = eta * mechwork + dmx * wu[4] + dmy * wv[4] + dmz * ww[4];
            const int64_t fo = iidx(g, gi - NG, gj - NG, k);
            const int64_t c = g.c;
            difflux[fo] = 0.0E0;
            difflux[fo + IMX * c] = dmx;
            difflux[fo + IMY * c] = dmy;
            difflux[fo + IMZ * c] = dmz;
            difflux[fo + IENE * c] = alam * (txx + tyy + tzz) + mechwork;
#undef XS
#undef YS
#undef ZW
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
#undef D1
#undef D2

constexpr int kHypBY = 4;
constexpr int kDiffBY = 4;

int pick_kz(const Geom &g, int tiles_xy)
{
    // Enough workgroups to fill 256 CUs several times over, with runs of planes
    // long enough that the 8-plane window warm-up stays a small overhead.
    int kz = g.n2;
    while (kz > 16 && (int64_t)tiles_xy * ((g.n2 + kz - 1) / kz) < 2048) kz = (kz + 1) / 2;
    return kz;
}

}  // namespace

hipError_t launch_hypterm_tiled(const Geom &g, const double *cons, const double *q, double *flux,
                                double dxi0, double dxi1, double dxi2, hipStream_t s)
{
    constexpr int BY = kHypBY;
    const int ntx = (g.n0 + TX - 1) / TX, nty = (g.n1 + BY - 1) / BY;
    const int kz = pick_kz(g, ntx * nty);
    const int ntz = (g.n2 + kz - 1) / kz;
    hypterm_tiled_kernel<BY><<<ntx * nty * ntz, TX * BY, 0, s>>>(g, cons, q, flux, dxi0, dxi1, dxi2,
                                                                kz, ntx, nty);
    return hipGetLastError();
}

hipError_t launch_diffterm_tiled(const Geom &g, const double *q, double *difflux, double dxi0,
                                 double dxi1, double dxi2, double eta, double alam, hipStream_t s)
{
    constexpr int BY = kDiffBY;
    const int ntx = (g.n0 + TX - 1) / TX, nty = (g.n1 + BY - 1) / BY;
    const int kz = pick_kz(g, ntx * nty);
    const int ntz = (g.n2 + kz - 1) / kz;
    diffterm_tiled_kernel<BY><<<ntx * nty * ntz, TX * BY, 0, s>>>(g, q, difflux, dxi0, dxi1, dxi2,
                                                                 eta, alam, kz, ntx, nty);
    return hipGetLastError();
}

}  // namespace cns
