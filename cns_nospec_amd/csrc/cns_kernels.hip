// cns_kernels.hip -- streaming kernels of the CNS_NoSpec step (gfx950):
// periodic ghost fill, ctoprim with the Courant max-reduction, the RK3 stage
// update, and the reference-order direct-load stencils used to cross-check the
// tiled ones.
//
// Every kernel keeps the reference's per-cell operation order; the library is
// compiled with -ffp-contract=off, so results are bit-identical to the
// reference cpu/ path (which g++ -O2 builds without FMA).
#include "cns_kernels.h"

namespace cns {

namespace {

constexpr int kStreamBlock = 256;

inline int stream_grid(int64_t work, int per_thread = 4)
{
    int64_t blocks = (work + (int64_t)kStreamBlock * per_thread - 1) / ((int64_t)kStreamBlock * per_thread);
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    return (int)blocks;
}

__device__ __forceinline__ int wrap_ghost(int a, int n)
{
    return a < NG ? a + n : (a >= n + NG ? a - n : a);
}

// ------------------------------------------------------------------ ghost fill
// The reference fills the ghost shell in three ordered copy passes (z planes over
// interior i,j; y planes over interior i and all k; x columns over all j,k:
// cpu/helper_functions.cpp:226-262, gpu/util.cu:53-127).  For extents >= NG their
// composition maps every ghost cell to the interior cell at the periodic image
// of all three of its indices, so one pass over the shell copies exactly the same
// values.  The shell is enumerated compactly: first the full rows whose j or k is
// a ghost index (g0 cells each), then the 2*NG ghost cells of each interior row.
__global__ void __launch_bounds__(kStreamBlock)
fill_boundary_kernel(Geom g, double *__restrict__ u, double *__restrict__ q, unsigned qmask,
                     int64_t n_rows_a, int64_t total)
{
    const int64_t plane_rows = (int64_t)2 * NG * g.g1;
    const int64_t elems_a = n_rows_a * g.g0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (e < elems_a) {
            int64_t r = e / g.g0;
            i = (int)(e - r * g.g0);
            if (r < plane_rows) {
                int kk = (int)(r / g.g1);
                j = (int)(r - (int64_t)kk * g.g1);
                k = kk < NG ? kk : g.n2 + kk;
            } else {
                r -= plane_rows;
                int kk = (int)(r / (2 * NG));
                int jj = (int)(r - (int64_t)kk * 2 * NG);
                k = NG + kk;
                j = jj < NG ? jj : g.n1 + jj;
            }
        } else {
            int64_t e2 = e - elems_a;
            int64_t row = e2 / (2 * NG);
            int ii = (int)(e2 - row * 2 * NG);
            i = ii < NG ? ii : g.n0 + ii;
            int kk = (int)(row / g.n1);
            k = NG + kk;
            j = NG + (int)(row - (int64_t)kk * g.n1);
        }
        const int64_t dst = gidx(g, i, j, k);
        const int64_t src = gidx(g, wrap_ghost(i, g.n0), wrap_ghost(j, g.n1), wrap_ghost(k, g.n2));
#pragma unroll
        for (int l = 0; l < NC; l++) u[l * g.cg + dst] = u[l * g.cg + src];
#pragma unroll
        for (int l = 0; l < NQ; l++)
            if (qmask & (1u << l)) q[l * g.cg + dst] = q[l * g.cg + src];
    }
}

// Ghost cell h of the compact shell enumeration of fill_boundary_kernel, as (i, j, k).
__device__ __forceinline__ void shell_cell(const Geom &g, int64_t e, int64_t elems_a, int &i, int &j, int &k)
{
    const int64_t plane_rows = (int64_t)2 * NG * g.g1;
    if (e < elems_a) {
        int64_t r = e / g.g0;
        i = (int)(e - r * g.g0);
        if (r < plane_rows) {
            const int kk = (int)(r / g.g1);
            j = (int)(r - (int64_t)kk * g.g1);
            k = kk < NG ? kk : g.n2 + kk;
        } else {
            r -= plane_rows;
            const int kk = (int)(r / (2 * NG));
            const int jj = (int)(r - (int64_t)kk * 2 * NG);
            k = NG + kk;
            j = jj < NG ? jj : g.n1 + jj;
        }
    } else {
        const int64_t e2 = e - elems_a;
        const int64_t row = e2 / (2 * NG);
        const int ii = (int)(e2 - row * 2 * NG);
        i = ii < NG ? ii : g.n0 + ii;
        const int kk = (int)(row / g.n1);
        k = NG + kk;
        j = NG + (int)(row - (int64_t)kk * g.n1);
    }
}

// One axis of the domain fill: ghost index a (0 <= a < n + 2 NG) of box I of nb
// boxes of n cells -> the box I' and ghosted index a' holding the periodic image.
__device__ __forceinline__ void domain_image(int a, int I, int n, int nb, int &Is, int &as)
{
    const int N = n * nb;
    int G = I * n + a - NG;  // global cell index, in [-NG, N + NG)
    G = G < 0 ? G + N : (G >= N ? G - N : G);
    Is = G / n;
    as = G - Is * n + NG;
}

// Domain fill: one y-row of the grid per box (blockIdx.y), grid-stride over its shell.
__global__ void __launch_bounds__(kStreamBlock)
domain_fill_kernel(Geom g, BoxSet bs, int nbx, int nby, int nbz, double *__restrict__ u,
                   double *__restrict__ q, unsigned qmask, int64_t n_rows_a, int64_t total)
{
    const int b = blockIdx.y;
    const int I = b % nbx, J = (b / nbx) % nby, K = b / (nbx * nby);
    const int64_t elems_a = n_rows_a * g.g0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        shell_cell(g, e, elems_a, i, j, k);
        int Is, Js, Ks, is, js, ks;
        domain_image(i, I, g.n0, nbx, Is, is);
        domain_image(j, J, g.n1, nby, Js, js);
        domain_image(k, K, g.n2, nbz, Ks, ks);
        const int bsrc = Is + nbx * (Js + nby * Ks);
        const int64_t dst = gidx(g, i, j, k), src = gidx(g, is, js, ks);
        double *ud = u + b * bs.su;
        const double *us = u + bsrc * bs.su;
#pragma unroll
        for (int l = 0; l < NC; l++) ud[l * g.cg + dst] = us[l * g.cg + src];
        if (qmask) {
            double *qd = q + b * bs.sq;
            const double *qs = q + bsrc * bs.sq;
#pragma unroll
            for (int l = 0; l < NQ; l++)
                if (qmask & (1u << l)) qd[l * g.cg + dst] = qs[l * g.cg + src];
        }
    }
}

// -------------------------------------------------------------------- courno
__global__ void courno_reset_kernel(unsigned long long *acc, double init)
{
    acc[0] = init > 0.0 ? (unsigned long long)__double_as_longlong(init) : 0ull;
}

// courno = MAX(MAX(courmx,courmy), MAX(courmz, courno_in)) (cpu/ctoprim.cpp:69) with
// courno_in = 1e-50 (cpu/advance.cpp:52); dt = cfl/courno (cpu/advance.cpp:57-58).
__global__ void courno_to_dt_kernel(const unsigned long long *acc, double cfl, double *dt,
                                    double *courno_out)
{
    double c = __longlong_as_double((long long)acc[0]);
    double courno = c > 1.0E-50 ? c : 1.0E-50;
    if (!(c == c)) courno = c;
    dt[0] = cfl / courno;
    if (courno_out) courno_out[0] = courno;
}

// ------------------------------------------------------------------- ctoprim
// Per ghosted cell (cpu/ctoprim.cpp:27-44):
//   rhoinv = 1/rho; q = (rho, mx*rhoinv, my*rhoinv, mz*rhoinv, (g-1)*eint*rho, eint/CV),
//   eint = E*rhoinv - 0.5*(u^2+v^2+w^2).
// With COURNO, per interior cell (cpu/ctoprim.cpp:48-63): c = sqrt(g*p/rho),
// cour_d = (c+|u_d|)/dx_d; the block's maximum goes to acc with one atomic.
// The grid strides over the padded linear index space, so one wave touches 64
// consecutive doubles of a row (coalesced); lanes on padding skip.
template <bool COURNO>
__global__ void __launch_bounds__(kStreamBlock)
ctoprim_kernel(Geom g, const double *__restrict__ u, double *__restrict__ q,
               unsigned long long *__restrict__ acc, double dx0, double dx1, double dx2, bool full)
{
    const int64_t total = (int64_t)g.g2 * g.pg;
    const int64_t cg = g.cg;
    unsigned long long cmax = 0ull;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / g.xg;
        const int i = (int)(e - row * g.xg);
        if (i >= g.g0) continue;
        const double rho = u[e];
        double qu, qv, qw, eint, p;
        ctoprim_cell(rho, u[e + cg], u[e + 2 * cg], u[e + 3 * cg], u[e + 4 * cg], qu, qv, qw, eint, p);
        if (full) q[e] = rho;
        q[e + cg] = qu;
        q[e + 2 * cg] = qv;
        q[e + 3 * cg] = qw;
        if (full) q[e + 4 * cg] = p;
        q[e + 5 * cg] = eint / CV;
        if (COURNO) {
            const int k = (int)(row / g.g1);
            const int j = (int)(row - (int64_t)k * g.g1);
            if (i >= NG && i < g.n0 + NG && j >= NG && j < g.n1 + NG && k >= NG && k < g.n2 + NG) {
                const double c = sqrt(GAMMA * p / rho);
                const double courx = (c + fabs(qu)) / dx0;
                const double coury = (c + fabs(qv)) / dx1;
                const double courz = (c + fabs(qw)) / dx2;
                unsigned long long b = courno_bits(courx);
                unsigned long long by = courno_bits(coury);
                unsigned long long bz = courno_bits(courz);
                b = b > by ? b : by;
                b = b > bz ? b : bz;
                cmax = cmax > b ? cmax : b;
            }
        }
    }
    if (COURNO) {
        __shared__ unsigned long long wmax[kStreamBlock / 64];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(cmax, off);
            cmax = cmax > o ? cmax : o;
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) wmax[wave] = cmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = wmax[0];
            for (int w = 1; w < kStreamBlock / 64; w++) m = m > wmax[w] ? m : wmax[w];
            atomicMax(acc, m);
        }
    }
}

// ------------------------------------------------------------------ RK update
// cpu/advance.cpp:78-85 (phase 1), :110-119 (phase 2), :144-153 (phase 3).
__global__ void __launch_bounds__(kStreamBlock)
rk_update_kernel(Geom g, int phase, double *__restrict__ u, double *__restrict__ unew,
                 const double *__restrict__ d, const double *__restrict__ f,
                 const double *__restrict__ dt_dev, double dt_host)
{
    const double dt = dt_dev ? dt_dev[0] : dt_host;
    const int64_t total = (int64_t)g.n2 * g.p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / g.x;
        const int i = (int)(e - row * g.x);
        if (i >= g.n0) continue;
        const int k = (int)(row / g.n1);
        const int j = (int)(row - (int64_t)k * g.n1);
        const int64_t o = gidx(g, i + NG, j + NG, k + NG);
#pragma unroll
        for (int l = 0; l < NC; l++) {
            const int64_t go = l * g.cg + o, io = l * g.c + e;
            if (phase == 1)
                unew[go] = u[go] + dt * (d[io] + f[io]);
            else if (phase == 2)
                unew[go] = ThreeQuarters * u[go] + OneQuarter * (unew[go] + dt * (d[io] + f[io]));
            else
                u[go] = OneThird * u[go] + TwoThirds * (unew[go] + dt * (d[io] + f[io]));
        }
    }
}

// ------------------------------------------------- direct-load stencils
__device__ __forceinline__ double d1(const double *__restrict__ f, int64_t o, int64_t s)
{
    return ALP * (f[o + s] - f[o - s]) + BET * (f[o + 2 * s] - f[o - 2 * s])
           + GAM * (f[o + 3 * s] - f[o - 3 * s]) + DEL * (f[o + 4 * s] - f[o - 4 * s]);
}

__device__ __forceinline__ double d2(const double *__restrict__ f, int64_t o, int64_t s)
{
    return CENTER * f[o] + OFF1 * (f[o + s] + f[o - s]) + OFF2 * (f[o + 2 * s] + f[o - 2 * s])
           + OFF3 * (f[o + 3 * s] + f[o - 3 * s]) + OFF4 * (f[o + 4 * s] + f[o - 4 * s]);
}

// Sums of one direction of cpu/hypterm.cpp (x :33-90, y :101-147, z :158-204) before *dxinv.
template <int DIR>
__device__ __forceinline__ void hyp_sums(const double *__restrict__ cons,
                                         const double *__restrict__ q, int64_t o, int64_t s,
                                         int64_t cg, double A[5])
{
    constexpr int qn = QU + DIR, mn = IMX + DIR;
    const double *cm = cons + mn * cg;
    const double *un = q + qn * cg;
    const double *pr = q + QPRES * cg;
#define UN(m) un[o + (m) * s]
#define P(m) pr[o + (m) * s]
#define CC(l, m) cons[(l) * cg + o + (m) * s]
    A[IRHO] = ALP * (cm[o + s] - cm[o - s]) + BET * (cm[o + 2 * s] - cm[o - 2 * s])
              + GAM * (cm[o + 3 * s] - cm[o - 3 * s]) + DEL * (cm[o + 4 * s] - cm[o - 4 * s]);
#pragma unroll
    for (int l = IMX; l <= IMZ; l++) {
        if (l == mn)
            A[l] = ALP * (CC(l, 1) * UN(1) - CC(l, -1) * UN(-1) + (P(1) - P(-1)))
                   + BET * (CC(l, 2) * UN(2) - CC(l, -2) * UN(-2) + (P(2) - P(-2)))
                   + GAM * (CC(l, 3) * UN(3) - CC(l, -3) * UN(-3) + (P(3) - P(-3)))
                   + DEL * (CC(l, 4) * UN(4) - CC(l, -4) * UN(-4) + (P(4) - P(-4)));
        else
            A[l] = ALP * (CC(l, 1) * UN(1) - CC(l, -1) * UN(-1))
                   + BET * (CC(l, 2) * UN(2) - CC(l, -2) * UN(-2))
                   + GAM * (CC(l, 3) * UN(3) - CC(l, -3) * UN(-3))
                   + DEL * (CC(l, 4) * UN(4) - CC(l, -4) * UN(-4));
    }
    A[IENE] = ALP * (CC(IENE, 1) * UN(1) - CC(IENE, -1) * UN(-1) + (P(1) * UN(1) - P(-1) * UN(-1)))
              + BET * (CC(IENE, 2) * UN(2) - CC(IENE, -2) * UN(-2) + (P(2) * UN(2) - P(-2) * UN(-2)))
              + GAM * (CC(IENE, 3) * UN(3) - CC(IENE, -3) * UN(-3) + (P(3) * UN(3) - P(-3) * UN(-3)))
              + DEL * (CC(IENE, 4) * UN(4) - CC(IENE, -4) * UN(-4) + (P(4) * UN(4) - P(-4) * UN(-4)));
#undef UN
#undef P
#undef CC
}

// cpu/hypterm.cpp:11-209, one thread per interior cell, x then y then z.
__global__ void __launch_bounds__(256)
hypterm_direct_kernel(Geom g, const double *__restrict__ cons, const double *__restrict__ q,
                      double *__restrict__ flux, double dxi0, double dxi1, double dxi2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.n0 || j >= g.n1) return;
    const int64_t o = gidx(g, i + NG, j + NG, k + NG);
    const int64_t fo = iidx(g, i, j, k);
    double A[5], F[5];
    hyp_sums<0>(cons, q, o, 1, g.cg, A);
#pragma unroll
    for (int l = 0; l < NC; l++) F[l] = -(A[l]) * dxi0;
    hyp_sums<1>(cons, q, o, g.xg, g.cg, A);
#pragma unroll
    for (int l = 0; l < NC; l++) F[l] = F[l] - (A[l]) * dxi1;
    hyp_sums<2>(cons, q, o, g.pg, g.cg, A);
#pragma unroll
    for (int l = 0; l < NC; l++) F[l] = F[l] - (A[l]) * dxi2;
#pragma unroll
    for (int l = 0; l < NC; l++) flux[l * g.c + fo] = F[l];
}

// cpu/diffterm.cpp:22-320, one thread per interior cell.  The first derivatives
// the reference stores in nine temporaries are recomputed at each use with the
// same expression, so each value is the same double.
__global__ void __launch_bounds__(256)
diffterm_direct_kernel(Geom g, const double *__restrict__ q, double *__restrict__ difflux,
                       double dxi0, double dxi1, double dxi2, double eta, double alam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.n0 || j >= g.n1) return;
    const int64_t o = gidx(g, i + NG, j + NG, k + NG);
    const int64_t fo = iidx(g, i, j, k);
    const int64_t sx = 1, sy = g.xg, sz = g.pg, c = g.c;
    const double *qU = q + QU * g.cg, *qV = q + QV * g.cg, *qW = q + QW * g.cg;
    const double *qT = q + QTEMP * g.cg;

#define UX(p) (d1(qU, (p), sx) * dxi0)
#define VY(p) (d1(qV, (p), sy) * dxi1)
#define WZ(p) (d1(qW, (p), sz) * dxi2)
#define D1OF(F, s, dxi) \
    ((ALP * (F(o + (s)) - F(o - (s))) + BET * (F(o + 2 * (s)) - F(o - 2 * (s))) \
      + GAM * (F(o + 3 * (s)) - F(o - 3 * (s))) + DEL * (F(o + 4 * (s)) - F(o - 4 * (s)))) * (dxi))

    const double uxx = d2(qU, o, sx) * (dxi0 * dxi0);
    const double uyy = d2(qU, o, sy) * (dxi1 * dxi1);
    const double uzz = d2(qU, o, sz) * (dxi2 * dxi2);
    const double vyx = D1OF(VY, sx, dxi0);
    const double wzx = D1OF(WZ, sx, dxi0);
    const double dmx = eta * (FourThirds * uxx + uyy + uzz + OneThird * (vyx + wzx));

    const double vxx = d2(qV, o, sx) * (dxi0 * dxi0);
    const double vyy = d2(qV, o, sy) * (dxi1 * dxi1);
    const double vzz = d2(qV, o, sz) * (dxi2 * dxi2);
    const double uxy = D1OF(UX, sy, dxi1);
    const double wzy = D1OF(WZ, sy, dxi1);
    const double dmy = eta * (vxx + FourThirds * vyy + vzz + OneThird * (uxy + wzy));

    const double wxx = d2(qW, o, sx) * (dxi0 * dxi0);
    const double wyy = d2(qW, o, sy) * (dxi1 * dxi1);
    const double wzz = d2(qW, o, sz) * (dxi2 * dxi2);
    const double uxz = D1OF(UX, sz, dxi2);
    const double vyz = D1OF(VY, sz, dxi2);
    const double dmz = eta * (wxx + wyy + FourThirds * wzz + OneThird * (uxz + vyz));

    const double txx = d2(qT, o, sx) * (dxi0 * dxi0);
    const double tyy = d2(qT, o, sy) * (dxi1 * dxi1);
    const double tzz = d2(qT, o, sz) * (dxi2 * dxi2);
    const double ux = UX(o), vy = VY(o), wz = WZ(o);
    const double uy = d1(qU, o, sy) * dxi1, uz = d1(qU, o, sz) * dxi2;
    const double vx = d1(qV, o, sx) * dxi0, vz = d1(qV, o, sz) * dxi2;
    const double wx = d1(qW, o, sx) * dxi0, wy = d1(qW, o, sy) * dxi1;
    const double divu = TwoThirds * (ux + vy + wz);
    const double tauxx = 2.E0 * ux - divu;
    const double tauyy = 2.E0 * vy - divu;
    const double tauzz = 2.E0 * wz - divu;
    const double tauxy = uy + vx;
    const double tauxz = uz + wx;
    const double tauyz = vz + wy;
    double mechwork = tauxx * ux + tauyy * vy + tauzz * wz + (tauxy * tauxy) + (tauxz * tauxz)
                      + (tauyz * tauyz);
    mechwork = eta * mechwork + dmx * qU[o] + dmy * qV[o] + dmz * qW[o];

    difflux[fo] = 0.0E0;
    difflux[fo + IMX * c] = dmx;
    difflux[fo + IMY * c] = dmy;
    difflux[fo + IMZ * c] = dmz;
    difflux[fo + IENE * c] = alam * (txx + tyy + tzz) + mechwork;
#undef UX
#undef VY
#undef WZ
#undef D1OF
}

}  // namespace

// ------------------------------------------------------------------ launchers
hipError_t launch_fill_boundary(const Geom &g, double *u, hipStream_t s)
{
    const int64_t rows_a = (int64_t)2 * NG * g.g1 + (int64_t)g.n2 * 2 * NG;
    const int64_t total = rows_a * g.g0 + (int64_t)g.n1 * g.n2 * 2 * NG;
    fill_boundary_kernel<<<stream_grid(total, 2), kStreamBlock, 0, s>>>(g, u, nullptr, 0u, rows_a, total);
    return hipGetLastError();
}

hipError_t launch_fill_boundary_q(const Geom &g, double *u, double *q, unsigned qmask, hipStream_t s)
{
    if (!q) qmask = 0;
    const int64_t rows_a = (int64_t)2 * NG * g.g1 + (int64_t)g.n2 * 2 * NG;
    const int64_t total = rows_a * g.g0 + (int64_t)g.n1 * g.n2 * 2 * NG;
    fill_boundary_kernel<<<stream_grid(total, 2), kStreamBlock, 0, s>>>(g, u, q, qmask & 63u, rows_a, total);
    return hipGetLastError();
}

hipError_t launch_domain_fill(const Geom &g, const BoxSet &bs, const int nb[3], double *u, double *q,
                              unsigned qmask, hipStream_t s)
{
    if (!q) qmask = 0;
    if (bs.n != nb[0] * nb[1] * nb[2] || bs.n < 1 || bs.n > 65535) return hipErrorInvalidValue;
    const int64_t rows_a = (int64_t)2 * NG * g.g1 + (int64_t)g.n2 * 2 * NG;
    const int64_t total = rows_a * g.g0 + (int64_t)g.n1 * g.n2 * 2 * NG;
    int gx = stream_grid(total, 2);
    // a few waves per box keep the whole grid resident on small boxes
    const int64_t want = (int64_t)(gx + bs.n - 1) / bs.n;
    gx = (int)(want < 1 ? 1 : want);
    const dim3 grid((unsigned)gx, (unsigned)bs.n);
    domain_fill_kernel<<<grid, kStreamBlock, 0, s>>>(g, bs, nb[0], nb[1], nb[2], u, q, qmask & 63u, rows_a, total);
    return hipGetLastError();
}

hipError_t launch_courno_reset(unsigned long long *acc, double init, hipStream_t s)
{
    courno_reset_kernel<<<1, 1, 0, s>>>(acc, init);
    return hipGetLastError();
}

hipError_t launch_ctoprim(const Geom &g, const double *u, double *q, unsigned long long *acc,
                          double dx0, double dx1, double dx2, hipStream_t s, bool full)
{
    const int64_t total = (int64_t)g.g2 * g.pg;
    const int grid = stream_grid(total, 8);
    if (acc)
        ctoprim_kernel<true><<<grid, kStreamBlock, 0, s>>>(g, u, q, acc, dx0, dx1, dx2, full);
    else
        ctoprim_kernel<false><<<grid, kStreamBlock, 0, s>>>(g, u, q, nullptr, dx0, dx1, dx2, full);
    return hipGetLastError();
}

hipError_t launch_courno_to_dt(const unsigned long long *acc, double cfl, double *dt_dev,
                               double *courno_dev, hipStream_t s)
{
    courno_to_dt_kernel<<<1, 1, 0, s>>>(acc, cfl, dt_dev, courno_dev);
    return hipGetLastError();
}

hipError_t launch_rk_update(const Geom &g, int phase, double *u, double *unew, const double *d,
                            const double *f, const double *dt_dev, double dt_host, hipStream_t s)
{
    const int64_t total = (int64_t)g.n2 * g.p;
    rk_update_kernel<<<stream_grid(total, 4), kStreamBlock, 0, s>>>(g, phase, u, unew, d, f, dt_dev,
                                                                     dt_host);
    return hipGetLastError();
}

hipError_t launch_hypterm_direct(const Geom &g, const double *cons, const double *q, double *flux,
                                 double dxi0, double dxi1, double dxi2, hipStream_t s)
{
    dim3 block(64, 4, 1), grid((g.n0 + 63) / 64, (g.n1 + 3) / 4, g.n2);
    hypterm_direct_kernel<<<grid, block, 0, s>>>(g, cons, q, flux, dxi0, dxi1, dxi2);
    return hipGetLastError();
}

hipError_t launch_diffterm_direct(const Geom &g, const double *q, double *difflux, double dxi0,
                                  double dxi1, double dxi2, double eta, double alam, hipStream_t s)
{
    dim3 block(64, 4, 1), grid((g.n0 + 63) / 64, (g.n1 + 3) / 4, g.n2);
    diffterm_direct_kernel<<<grid, block, 0, s>>>(g, q, difflux, dxi0, dxi1, dxi2, eta, alam);
    return hipGetLastError();
}

}  // namespace cns
