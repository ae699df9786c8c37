// cns_stencil_pipe.hip -- pipelined z-marching hypterm and diffterm for gfx950.
//
// A workgroup owns a TX x 8 column tile (TX = 64: 512 threads, one wave per row; the
// step's fused kernels use TX = 32: 256 threads, two rows per wave, two workgroups per
// CU) and marches up a run of planes.  Every field whose x/y neighbours a plane needs
// is staged in LDS as a full 16 x (TX+8) tile (4 cells of halo on each side), loaded with
// direct-to-LDS DMA (global_load_lds_dwordx4: each lane moves 16 bytes straight
// into the tile, no registers involved) into one of two LDS buffers.  The loads of
// plane k+1 are issued right after the barrier that opens plane k, so the whole
// compute of plane k hides their latency; a plane costs one barrier (hypterm) or
// two (diffterm), and there is no per-plane gather loop.  z neighbours of the
// lane's own column come from 9-plane register windows whose leading values are
// prefetched one plane ahead.
//
// diffterm reads wz = d(w)/dz on the cross halo as a precomputed field (wz_kernel
// below, the reference's wz temporary of cpu/diffterm.cpp:127-151 over the ghosted
// x-y extent), so no plane needs the 8 z neighbours of halo columns.
//
// Arithmetic and its order are the reference's (cpu/hypterm.cpp:11-209,
// cpu/diffterm.cpp:22-320, cpu/advance.cpp:78-153); with -ffp-contract=off the
// results are bit-identical to the reference cpu/ path and to the other kernels.
#include <cstdlib>

#include "cns_kernels.h"

namespace cns {

namespace {

typedef __attribute__((address_space(3))) void lds_void;

#define D1(F) (ALP * (F(1) - F(-1)) + BET * (F(2) - F(-2)) + GAM * (F(3) - F(-3)) + DEL * (F(4) - F(-4)))
#define D2(F)                                                                                \
    (CENTER * F(0) + OFF1 * (F(1) + F(-1)) + OFF2 * (F(2) + F(-2)) + OFF3 * (F(3) + F(-3)) \
     + OFF4 * (F(4) + F(-4)))
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

__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int xcd = b & 7, local = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Tile geometry shared by the DMA issue: the ghosted-array coordinates of tile cell
// (0, 0) and the clamps that keep every lane's 16-byte load inside the array (cells
// beyond the box are only read by lanes whose outputs are discarded).
struct TileDma {
    int gi0, gj0, g1m1, xgm2;
    unsigned xg;
};

// All of this wave's loads, DMA included, have landed; then the workgroup barrier.
__device__ __forceinline__ void plane_barrier()
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Tile shapes of the pipelined stencils: TXs x 8 cells per workgroup of TXs*8
// threads.  At 64 x 8 diffterm takes 142 KB of LDS (one workgroup per CU); at 32 x 8
// 77 KB, so two workgroups share a CU and one's barrier waits are the other's compute
// time (likewise the fused hypterm: 117 KB and 65 KB).
template <int TXs>
struct Shape {
    static constexpr int TX = TXs, BY = 8, PX = TX + 8, R = BY + 8, NT = TX * BY, NW = NT / 64;
    static constexpr int TILE = R * PX, NDMA = TILE / 128;
    static_assert(TILE % 128 == 0, "a tile is a whole number of 1 KB DMA blocks");
    static_assert(PX % 2 == 0, "a lane's 16-byte pair stays in one row");
};

// DMA block n of one field tile: lane L moves tile elements e, e+1 with e = 128 n + 2 L
// (row e / PX, column e % PX; PX is even, so a pair never straddles rows) from the
// plane at `src` into tile + 128 n + 2 L.  A plane offset fits 32 bits; keeping it
// 32-bit halves what the compiler holds live across the plane loop.
template <class S>
__device__ __forceinline__ void dma_block_s(const double *src, double *tile, int n, const TileDma &t, int lane)
{
    const int e = 128 * n + 2 * lane;
    const int r = e / S::PX, c = e - r * S::PX;
    const int hj = min(t.gj0 + r, t.g1m1), hi = min(t.gi0 + c, t.xgm2);
    const unsigned off = (unsigned)hj * t.xg + (unsigned)hi;
    __builtin_amdgcn_global_load_lds((const void *)(src + off), (lds_void *)(tile + 128 * n), 16, 0, 0);
}

// This wave's share of the NF x NDMA DMA blocks of one plane: block n of field f goes
// to wave (n + f) mod NW (NW waves; the last NDMA - NW blocks to the waves after them).
template <class S, int NF>
__device__ __forceinline__ void dma_plane_s(const double *const (&src)[NF], int64_t plane, double *buf,
                                            const TileDma &t, int wave, int lane)
{
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const int n0 = (wave - f) & (S::NW - 1);
        dma_block_s<S>(src[f] + plane, buf + f * S::TILE, n0, t, lane);
        if (n0 + S::NW < S::NDMA) dma_block_s<S>(src[f] + plane, buf + f * S::TILE, n0 + S::NW, t, lane);
    }
}

struct PipeRk {
    const double *d;     // difflux (interior)
    const double *base;  // U^n (ghosted), phases 2 and 3
    double *out;         // result (ghosted); must not alias cons
    const double *dt;    // device dt
    int phase;           // 0: plain hypterm
    double *qout;        // primitives of the result (interior cells), or null
};

// hypterm's flux of the lane's cell (LDS index me of 16 x P tiles): x terms from the
// row, y terms from the column, z terms from the register windows (plane k+m at
// [4+m]), in the reference's order (cpu/hypterm.cpp:33-90, :101-147, :158-204).
template <int P>
__device__ __forceinline__ void hyp_flux(const double *s_mx, const double *s_my, const double *s_mz,
                                         const double *s_E, const double *s_p, const double *s_u,
                                         const double *s_v, int me, const double (&wmx)[9],
                                         const double (&wmy)[9], const double (&wmz)[9],
                                         const double (&wE)[9], const double (&ww)[9],
                                         const double (&wp)[9], double dxi0, double dxi1,
                                         double dxi2, double (&F)[5])
{
    // ---- x: F = -(sum)*dxinv(1)
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
    // ---- y: F = F - (sum)*dxinv(2)
#define YMX(m) s_mx[me + (m) * P]
#define YMY(m) s_my[me + (m) * P]
#define YMZ(m) s_mz[me + (m) * P]
#define YE(m) s_E[me + (m) * P]
#define YP(m) s_p[me + (m) * P]
#define YV(m) s_v[me + (m) * P]
    F[IRHO] = F[IRHO] - (D1(YMY)) * dxi1;
    F[IMX] = F[IMX] - (HMOM(YMX, YV)) * dxi1;
    F[IMY] = F[IMY] - (HMOMP(YMY, YV, YP)) * dxi1;
    F[IMZ] = F[IMZ] - (HMOM(YMZ, YV)) * dxi1;
    F[IENE] = F[IENE] - (HENE(YE, YV, YP)) * dxi1;
    // ---- z: from the register windows
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
#undef XMX
#undef XMY
#undef XMZ
#undef XE
#undef XP
#undef XU
#undef YMX
#undef YMY
#undef YMZ
#undef YE
#undef YP
#undef YV
#undef ZMX
#undef ZMY
#undef ZMZ
#undef ZE
#undef ZW
#undef ZP
}

// u, v, w and p of one cell from its conserved values: ctoprim's arithmetic
// (cpu/ctoprim.cpp:80-100; the same expressions as ctoprim_kernel), so the values are
// bit-identical to what ctoprim stores in q.
__device__ __forceinline__ void prim_uvwp(double rho, double mx, double my, double mz, double E,
                                          double &u, double &v, double &w, double &p)
{
    double eint;
    ctoprim_cell(rho, mx, my, mz, E, u, v, w, eint, p);
}

// The stage update of the cell and its stores (STAGE1: cons + dt(D+F), the reference's
// form of cpu/advance.cpp:82; later stages ra*base + rb*(cons + dt(D+F)), which is
// 3/4 base + 1/4(..) and 1/3 base + 2/3(..) of :114-116, :148-150).
template <bool STAGE1, bool WRITE_F>
__device__ __forceinline__ void rk_store(const double (&F)[5], const double (&uc)[5], const double (&rkd)[5],
                                         const double (&rkb)[5], double ra, double rb, double dt,
                                         double *flux, int64_t fo, int64_t c, double *out, int64_t zo,
                                         int64_t cg, double (&o)[5])
{
    if (WRITE_F) {
#pragma unroll
        for (int l = 0; l < NC; l++) flux[l * c + fo] = F[l];
    }
#pragma unroll
    for (int l = 0; l < NC; l++) {
        o[l] = STAGE1 ? uc[l] + dt * (rkd[l] + F[l]) : ra * rkb[l] + rb * (uc[l] + dt * (rkd[l] + F[l]));
        out[l * cg + zo] = o[l];
    }
}

// ctoprim of the stage result o at interior cell zo (cpu/ctoprim.cpp:23-38), stored
// into q: u, v, w and T (QOUT 1: what the next stage's diffterm reads) or every
// component (QOUT 2: the Q the reference leaves after a step).  The ghost cells are
// the periodic fill's, so q over the ghosted box equals ctoprim of the filled result.
template <int QOUT>
__device__ __forceinline__ void q_store(const double (&o)[5], double *q, int64_t zo, int64_t cg)
{
    if (QOUT == 0) return;
    double u, v, w, eint, p;
    ctoprim_cell(o[IRHO], o[IMX], o[IMY], o[IMZ], o[IENE], u, v, w, eint, p);
    if (QOUT == 2) q[QRHO * cg + zo] = o[IRHO];
    q[QU * cg + zo] = u;
    q[QV * cg + zo] = v;
    q[QW * cg + zo] = w;
    if (QOUT == 2) q[QPRES * cg + zo] = p;
    q[QTEMP * cg + zo] = eint / CV;
}

// ======================================================================= hypterm
// LDS fields (centre plane): mx, my, mz, E, p, u, v.  Register windows: mx, my, mz,
// E, w, p.  With RK the stage update of the cell is applied as in cns_stencil.hip,
// with three savings of HBM traffic that leave every result bit-identical:
//  * D_rho is not read: diffterm always stores 0.0 there (cpu/diffterm.cpp:63-68);
//  * STAGE1: the base is cons itself, so the update is cons + dt*(D+F), the
//    reference's own form (cpu/advance.cpp:82), without re-reading the base;
//  * !WRITE_F: flux is not stored (stages 1 and 2 of a step, whose F the third
//    stage overwrites).
enum { HMX = 0, HMY, HMZ, HE, HP, HU, HV, HNF };

template <class S, bool RK, bool STAGE1, bool WRITE_F>
__global__ void __launch_bounds__(S::NT, 2)
hypterm_pipe_kernel(Geom g, const double *__restrict__ cons, const double *__restrict__ q,
                    double *__restrict__ flux, double dxi0, double dxi1, double dxi2, int kz,
                    int ntx, int nty, PipeRk rk)
{
    constexpr int TILE_ = S::TILE, PX_ = S::PX;
    __shared__ double sm[2 * HNF * TILE_];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;       // DMA issue
    const int col = threadIdx.x % S::TX, row = threadIdx.x / S::TX;  // the lane's cell
    const TileDma td{bx * S::TX, by * S::BY, g.g1 - 1, g.xg - 2, (unsigned)g.xg};
    const int gi = td.gi0 + col + 4, gj = td.gj0 + row + 4;
    const bool inbox = gi < g.g0 && gj < g.g1;
    const bool active = gi < g.n0 + NG && gj < g.n1 + NG;
    const int kbeg = bz * kz, kend = min(kbeg + kz, g.n2);
    const int64_t cg = g.cg, pg = g.pg;
    const double *c_mx = cons + IMX * cg, *c_my = cons + IMY * cg, *c_mz = cons + IMZ * cg,
                 *c_E = cons + IENE * cg;
    const double *q_u = q + QU * cg, *q_v = q + QV * cg, *q_w = q + QW * cg, *q_p = q + QPRES * cg;
    const double *const src[HNF] = {c_mx, c_my, c_mz, c_E, q_p, q_u, q_v};
    const int64_t own = (int64_t)gj * g.xg + gi;
    const int me = (row + 4) * PX_ + col + 4;
    double dt = 0.0, ra = 0.0, rb = 1.0;
    if (RK) {
        // r = ra*base + rb*(cons + dt*(D+F)) with (ra, rb) = (0, 1), (3/4, 1/4), (1/3, 2/3):
        // 0*base + 1*x is exactly x, so each phase keeps the reference's rounding
        // (cpu/advance.cpp:82, :114-116, :148-150)
        dt = rk.dt[0];
        ra = rk.phase == 1 ? 0.0E0 : (rk.phase == 2 ? ThreeQuarters : OneThird);
        rb = rk.phase == 1 ? 1.0E0 : (rk.phase == 2 ? OneQuarter : TwoThirds);
    }

    // first plane's tiles in flight while the windows fill
    dma_plane_s<S, HNF>(src, (int64_t)(kbeg + NG) * pg, sm, td, wave, lane);

    double wmx[9], wmy[9], wmz[9], wE[9], ww[9], wp[9];
    double nmx = 0, nmy = 0, nmz = 0, nE = 0, nw = 0, np = 0;
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wmx[m] = c_mx[o]; wmy[m] = c_my[o]; wmz[m] = c_mz[o];
            wE[m] = c_E[o]; ww[m] = q_w[o]; wp[m] = q_p[o];
        }
        const int64_t o = (int64_t)(kbeg + 2 * NG) * pg + own;
        nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o]; nw = q_w[o]; np = q_p[o];
    }

    for (int k = kbeg; k < kend; k++) {
        const int cur = (k - kbeg) & 1;
        const double *b = sm + cur * (HNF * TILE_);
        plane_barrier();
        if (k + 1 < kend)
            dma_plane_s<S, HNF>(src, (int64_t)(k + 1 + NG) * pg, sm + (cur ^ 1) * (HNF * TILE_), td, wave, lane);

        const int64_t zo = (int64_t)(k + NG) * pg + own;
        wmx[8] = nmx; wmy[8] = nmy; wmz[8] = nmz; wE[8] = nE; ww[8] = nw; wp[8] = np;
        if (inbox && k + 1 < kend) {
            const int64_t o = zo + 5 * pg;
            nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o]; nw = q_w[o]; np = q_p[o];
        }
        const int64_t fo = iidx(g, gi - NG, gj - NG, k);
        double rkd[5], rkb[5], rho_c = 0.0;
        if (RK && active) {
            rho_c = cons[zo];
            rkd[IRHO] = 0.0E0;
#pragma unroll
            for (int l = 0; l < NC; l++) {
                if (l != IRHO) rkd[l] = rk.d[l * g.c + fo];
                if (!STAGE1) rkb[l] = rk.base[l * cg + zo];
            }
        }

        if (active) {
            const double *s_mx = b + HMX * TILE_, *s_my = b + HMY * TILE_, *s_mz = b + HMZ * TILE_,
                         *s_E = b + HE * TILE_, *s_p = b + HP * TILE_, *s_u = b + HU * TILE_,
                         *s_v = b + HV * TILE_;
            double F[5];
            hyp_flux<PX_>(s_mx, s_my, s_mz, s_E, s_p, s_u, s_v, me, wmx, wmy, wmz, wE, ww, wp, dxi0, dxi1, dxi2, F);
            if (RK) {
                const double uc[5] = {rho_c, wmx[4], wmy[4], wmz[4], wE[4]};
                double o[5];
                rk_store<STAGE1, WRITE_F>(F, uc, rkd, rkb, ra, rb, dt, flux, fo, g.c, rk.out, zo, cg, o);
            } else {
#pragma unroll
                for (int l = 0; l < NC; l++) flux[l * g.c + fo] = F[l];
            }
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1];
            wE[m] = wE[m + 1]; ww[m] = ww[m + 1]; wp[m] = wp[m + 1];
        }
    }
}

// ====================================================== hypterm from the conserved state
// The fused stage inside cns_advance, where q is ctoprim(cons) by construction: only
// rho, mx, my, mz, E are staged by DMA, and u, v, p of the tile's cross cells are
// derived into LDS with ctoprim's arithmetic after the plane barrier (a second
// barrier publishes them); the windows' w and p are derived from the leading cell's
// conserved values.  Three ghosted fields fewer to fetch per plane than the q-based
// kernel, for the same bits.
enum { CRHO = 0, CMX, CMY, CMZ, CE, CNF };
enum { SDU = 0, SDV, SDP, SDNF };

template <class S, bool STAGE1, bool WRITE_F, int QOUT>
__global__ void __launch_bounds__(S::NT, 2)
hypterm_cons_kernel(Geom g, const double *__restrict__ cons, double *__restrict__ flux, double dxi0,
                    double dxi1, double dxi2, int kz, int ntx, int nty, PipeRk rk, BoxSet bs, int nbpb)
{
    constexpr int TILE_ = S::TILE, PX_ = S::PX;
    constexpr int XSTRIP = S::BY * PX_;           // rows 4..11, all PX columns
    constexpr int CROSS = XSTRIP + 8 * S::TX;    // + rows 0..3 and 12..15, columns 4..TX+3
    __shared__ double sm[2 * CNF * TILE_];
    __shared__ double sd[SDNF * TILE_];

    // box of a box set (nbpb workgroups each), then the tile within it
    int t = xcd_tile(blockIdx.x, gridDim.x);
    const int box = t / nbpb;
    t -= box * nbpb;
    cons += box * bs.su;
    flux += box * bs.sc;
    rk.d += box * bs.sc;
    rk.base += box * bs.su;
    rk.out += box * bs.su;
    if (rk.qout) rk.qout += box * bs.sq;
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;       // DMA issue
    const int col = threadIdx.x % S::TX, row = threadIdx.x / S::TX;  // the lane's cell
    const TileDma td{bx * S::TX, by * S::BY, g.g1 - 1, g.xg - 2, (unsigned)g.xg};
    const int gi = td.gi0 + col + 4, gj = td.gj0 + row + 4;
    const bool inbox = gi < g.g0 && gj < g.g1;
    const bool active = gi < g.n0 + NG && gj < g.n1 + NG;
    const int kbeg = bz * kz, kend = min(kbeg + kz, g.n2);
    const int64_t cg = g.cg, pg = g.pg;
    const double *c_rho = cons, *c_mx = cons + IMX * cg, *c_my = cons + IMY * cg,
                 *c_mz = cons + IMZ * cg, *c_E = cons + IENE * cg;
    const double *const src[CNF] = {c_rho, c_mx, c_my, c_mz, c_E};
    const int64_t own = (int64_t)gj * g.xg + gi;
    const int me = (row + 4) * PX_ + col + 4;
    const double dt = rk.dt[0];
    const double ra = rk.phase == 1 ? 0.0E0 : (rk.phase == 2 ? ThreeQuarters : OneThird);
    const double rb = rk.phase == 1 ? 1.0E0 : (rk.phase == 2 ? OneQuarter : TwoThirds);

    dma_plane_s<S, CNF>(src, (int64_t)(kbeg + NG) * pg, sm, td, wave, lane);

    double wmx[9], wmy[9], wmz[9], wE[9], ww[9], wp[9];
    double nrho = 1.0, nmx = 0, nmy = 0, nmz = 0, nE = 0;
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wmx[m] = c_mx[o]; wmy[m] = c_my[o]; wmz[m] = c_mz[o]; wE[m] = c_E[o];
            double u_, v_;
            prim_uvwp(c_rho[o], wmx[m], wmy[m], wmz[m], wE[m], u_, v_, ww[m], wp[m]);
        }
        const int64_t o = (int64_t)(kbeg + 2 * NG) * pg + own;
        nrho = c_rho[o]; nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o];
    }

    for (int k = kbeg; k < kend; k++) {
        const int cur = (k - kbeg) & 1;
        const double *b = sm + cur * (CNF * TILE_);
        plane_barrier();
        if (k + 1 < kend)
            dma_plane_s<S, CNF>(src, (int64_t)(k + 1 + NG) * pg, sm + (cur ^ 1) * (CNF * TILE_), td, wave, lane);

        const int64_t zo = (int64_t)(k + NG) * pg + own;
        {
            wmx[8] = nmx; wmy[8] = nmy; wmz[8] = nmz; wE[8] = nE;
            double u_, v_;
            prim_uvwp(nrho, nmx, nmy, nmz, nE, u_, v_, ww[8], wp[8]);
        }
        if (inbox && k + 1 < kend) {
            const int64_t o = zo + 5 * pg;
            nrho = c_rho[o]; nmx = c_mx[o]; nmy = c_my[o]; nmz = c_mz[o]; nE = c_E[o];
        }
        const int64_t fo = iidx(g, gi - NG, gj - NG, k);
        double rkd[5], rkb[5];
        if (active) {
            rkd[IRHO] = 0.0E0;
#pragma unroll
            for (int l = 0; l < NC; l++) {
                if (l != IRHO) rkd[l] = rk.d[l * g.c + fo];
                if (!STAGE1) rkb[l] = rk.base[l * cg + zo];
            }
        }

        // u, v, p of the cross cells of this plane's tile
        for (int h = threadIdx.x; h < CROSS; h += S::NT) {
            int r, c;
            if (h < XSTRIP) {
                r = 4 + h / PX_;
                c = h - (r - 4) * PX_;
            } else {
                const int h2 = h - XSTRIP, rr = h2 / S::TX;
                r = rr < 4 ? rr : rr + S::BY;
                c = 4 + h2 % S::TX;
            }
            const int l = r * PX_ + c;
            double u_, v_, w_, p_;
            prim_uvwp(b[CRHO * TILE_ + l], b[CMX * TILE_ + l], b[CMY * TILE_ + l], b[CMZ * TILE_ + l],
                      b[CE * TILE_ + l], u_, v_, w_, p_);
            sd[SDU * TILE_ + l] = u_;
            sd[SDV * TILE_ + l] = v_;
            sd[SDP * TILE_ + l] = p_;
        }
        __syncthreads();

        if (active) {
            double F[5];
            hyp_flux<PX_>(b + CMX * TILE_, b + CMY * TILE_, b + CMZ * TILE_, b + CE * TILE_, sd + SDP * TILE_,
                          sd + SDU * TILE_, sd + SDV * TILE_, me, wmx, wmy, wmz, wE, ww, wp, dxi0, dxi1, dxi2, F);
            const double uc[5] = {b[CRHO * TILE_ + me], wmx[4], wmy[4], wmz[4], wE[4]};
            double o[5];
            rk_store<STAGE1, WRITE_F>(F, uc, rkd, rkb, ra, rb, dt, flux, fo, g.c, rk.out, zo, cg, o);
            q_store<QOUT>(o, rk.qout, zo, cg);
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            wmx[m] = wmx[m + 1]; wmy[m] = wmy[m + 1]; wmz[m] = wmz[m + 1];
            wE[m] = wE[m + 1]; ww[m] = ww[m + 1]; wp[m] = wp[m + 1];
        }
    }
}

// ========================================================================= ctoprim+wz
// ctoprim over the ghosted box (cpu/ctoprim.cpp:11-105), z-marching: one thread per
// (i, j) column of the ghosted plane and a run of kz planes.  The column's w window
// also gives wz = D1_z(w)*dxinv(3) for the run's interior planes (cpu/diffterm.cpp:
// 126-151, the field diffterm_pipe_kernel reads), so the step needs no separate wz
// pass; the 4 planes on either side of the run are read for w only.  COURNO: the
// Courant maximum over interior cells (cpu/ctoprim.cpp:42-69) into acc, one 64-bit
// atomicMax per block.  !FULL: only u, v, w and T are stored (what diffterm reads).
constexpr int kCtoBlock = 256;

template <bool COURNO, bool FULL>
__global__ void __launch_bounds__(kCtoBlock)
ctoprim_wz_kernel(Geom g, const double *__restrict__ u, double *__restrict__ q, double *__restrict__ wz,
                  unsigned long long *__restrict__ acc, double dx0, double dx1, double dx2, double dxi2,
                  int kz, BoxSet bs)
{
    u += blockIdx.z * bs.su;  // box of a box set
    q += blockIdx.z * bs.sq;
    wz += blockIdx.z * bs.sw;
    const int64_t pcol = (int64_t)blockIdx.x * kCtoBlock + threadIdx.x;  // j*xg + i
    const int i = (int)(pcol % g.xg), j = (int)(pcol / g.xg);
    const bool col = pcol < g.pg && i < g.g0;
    const int kbeg = blockIdx.y * kz, kend = min(kbeg + kz, g.g2);
    const int64_t cg = g.cg, pg = g.pg;
    const bool xy_interior = i >= NG && i < g.n0 + NG && j >= NG && j < g.n1 + NG;
    unsigned long long cmax = 0ull;
    double w[9];
#pragma unroll
    for (int m = 0; m < 9; m++) w[m] = 0.0;
    const int k0 = max(kbeg - NG, 0), k1 = min(kend + NG, g.g2);
    for (int kk = k0; kk < k1; kk++) {
        const int64_t e = (int64_t)kk * pg + pcol;
        double wv = 0.0;
        if (col) {
            if (kk >= kbeg && kk < kend) {
                const double rho = u[e];
                double qu, qv, qw, eint, p;
                ctoprim_cell(rho, u[e + cg], u[e + 2 * cg], u[e + 3 * cg], u[e + 4 * cg], qu, qv, qw, eint, p);
                if (FULL) q[e] = rho;
                q[e + cg] = qu;
                q[e + 2 * cg] = qv;
                q[e + 3 * cg] = qw;
                if (FULL) q[e + 4 * cg] = p;
                q[e + 5 * cg] = eint / CV;
                wv = qw;
                if (COURNO && xy_interior && kk >= NG && kk < g.n2 + NG) {
                    const double c = sqrt(GAMMA * p / rho);
                    unsigned long long b = courno_bits((c + fabs(qu)) / dx0);
                    const unsigned long long by = courno_bits((c + fabs(qv)) / dx1);
                    const unsigned long long bz = courno_bits((c + fabs(qw)) / dx2);
                    b = b > by ? b : by;
                    b = b > bz ? b : bz;
                    cmax = cmax > b ? cmax : b;
                }
            } else {
                // w of a plane beyond the run: ctoprim's w exactly
                double qu, qv, eint, p;
                ctoprim_cell(u[e], u[e + cg], u[e + 2 * cg], u[e + 3 * cg], u[e + 4 * cg], qu, qv, wv, eint, p);
            }
        }
#pragma unroll
        for (int m = 0; m < 8; m++) w[m] = w[m + 1];
        w[8] = wv;
        // the window now holds planes kk-8 .. kk; its centre c = kk-4 is complete when
        // c-4 >= k0, which holds for every interior c of the run
        const int c = kk - NG;
        if (col && c >= kbeg && c < kend && c >= NG && c < g.n2 + NG) {
#define WW(mm) w[4 + (mm)]
            wz[(int64_t)c * pg + pcol] = D1(WW) * dxi2;
#undef WW
        }
    }
    if (COURNO) {
        __shared__ unsigned long long wmax[kCtoBlock / 64];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(cmax, off);
            cmax = cmax > o ? cmax : o;
        }
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = cmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = wmax[0];
            for (int wv = 1; wv < kCtoBlock / 64; wv++) m = m > wmax[wv] ? m : wmax[wv];
            atomicMax(acc, m);
        }
    }
}

// ============================================================================ wz
// wz(i, j, k) = D1_z(w) * dxinv(3) for every (i, j) of the ghosted x-y extent and
// interior k (cpu/diffterm.cpp:126-151), into a one-component array with the ghosted
// layout (plane k+ng).  One thread per column and run of kz planes, 9-plane window.
__global__ void __launch_bounds__(256)
wz_kernel(Geom g, const double *__restrict__ q_w, double *__restrict__ wz, double dxi2, int kz, BoxSet bs)
{
    q_w += blockIdx.z * bs.sq;  // box of a box set
    wz += blockIdx.z * bs.sw;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;  // j*xg + i
    if (p >= g.pg || (int)(p % g.xg) >= g.g0) return;
    const int kbeg = blockIdx.y * kz, kend = min(kbeg + kz, g.n2);
    const int64_t pg = g.pg;
    double w[9];
#pragma unroll
    for (int m = 0; m < 8; m++) w[m] = q_w[(int64_t)(kbeg + m) * pg + p];
    for (int k = kbeg; k < kend; k++) {
        w[8] = q_w[(int64_t)(k + 2 * NG) * pg + p];
#define WZ(mm) w[4 + (mm)]
        wz[(int64_t)(k + NG) * pg + p] = D1(WZ) * dxi2;
#undef WZ
#pragma unroll
        for (int m = 0; m < 8; m++) w[m] = w[m + 1];
    }
}

// ====================================================================== diffterm
// LDS fields, double-buffered: centre plane u, v (square halos), w, T, wz (cross
// halos), leading plane (k+4) u and v.  Single-buffered: ux on the tile and y-halo
// rows, vy on the tile and x-halo columns.  Per plane:
//  B  (after the opening barrier) ux, vy of the leading plane into the register
//     windows; ux, vy of the centre plane on their halos into LDS;
//  C  (after a second barrier) every output of cpu/diffterm.cpp:155-315 for the cell.
// Register windows: u, v, w, T and the computed ux, vy.
enum { DCU = 0, DCV, DCW, DCT, DWZ, DLU, DLV, DNF };

template <class S, bool WRITE_RHO>
__global__ void __launch_bounds__(S::NT, 2)
diffterm_pipe_kernel(Geom g, const double *__restrict__ q, const double *__restrict__ wzf,
                     double *__restrict__ difflux, double dxi0, double dxi1, double dxi2, double eta,
                     double alam, int kz, int ntx, int nty, BoxSet bs, int nbpb)
{
    __shared__ double sm[2 * DNF * S::TILE];
    __shared__ double s_ux[S::R * S::TX];   // row r (0..15), column 4+c
    __shared__ double s_vy[S::BY * S::PX];  // row 4+r (r 0..7), column c (0..PX-1)

    int t = xcd_tile(blockIdx.x, gridDim.x);
    const int box = t / nbpb;
    t -= box * nbpb;
    q += box * bs.sq;
    wzf += box * bs.sw;
    difflux += box * bs.sc;
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;       // DMA issue
    const int col = threadIdx.x % S::TX, row = threadIdx.x / S::TX;  // the lane's cell
    const TileDma td{bx * S::TX, by * S::BY, g.g1 - 1, g.xg - 2, (unsigned)g.xg};
    const int gi = td.gi0 + col + 4, gj = td.gj0 + row + 4;
    const bool inbox = gi < g.g0 && gj < g.g1;
    const bool active = gi < g.n0 + NG && gj < g.n1 + NG;
    const int kbeg = bz * kz, kend = min(kbeg + kz, g.n2);
    const int64_t cg = g.cg, pg = g.pg, xg = g.xg;
    const double *q_u = q + QU * cg, *q_v = q + QV * cg, *q_w = q + QW * cg, *q_t = q + QTEMP * cg;
    const double *const src[DNF] = {q_u, q_v, q_w, q_t, wzf, q_u + 4 * pg, q_v + 4 * pg};
    const int64_t own = (int64_t)gj * xg + gi;
    const int me = (row + 4) * S::PX + col + 4;  // own cell in R x PX tiles
    const int me64 = (row + 4) * S::TX + col;    // own cell in s_ux
    const int meby = row * S::PX + col + 4;      // own cell in s_vy
    const double dxi0s = dxi0 * dxi0, dxi1s = dxi1 * dxi1, dxi2s = dxi2 * dxi2;

    dma_plane_s<S, DNF>(src, (int64_t)(kbeg + NG) * pg, sm, td, wave, lane);

    // windows: plane k+m at [4+m]; ux and vy of the first 8 planes straight from
    // global memory (the same expressions as the per-plane path)
    // The tile's own-row ux and own-cell vy feed phase B's LDS planes (uxy, vyx of the
    // neighbours), so ux is needed on every x-interior lane of the ghosted y range and
    // vy on every y-interior lane of the ghosted x range; both stencils stay in bounds.
    double wu[9], wv[9], ww[9], wt[9], wux[9], wvy[9];
    double nw = 0.0, nt = 0.0;
    const bool need_ux = gi < g.n0 + NG && gj < g.g1, need_vy = gj < g.n1 + NG && gi < g.g0;
#pragma unroll 1
    for (int m = 0; m < 8; m++) {
        const int64_t o = (int64_t)(kbeg + m) * pg + own;
        double ux_ = 0.0, vy_ = 0.0;
#define GUX(mm) q_u[o + (mm)]
#define GVY(mm) q_v[o + (int64_t)(mm) * xg]
        if (need_ux) ux_ = D1(GUX) * dxi0;
        if (need_vy) vy_ = D1(GVY) * dxi1;
#undef GUX
#undef GVY
#pragma unroll
        for (int mm = 0; mm < 7; mm++) {
            wux[mm] = wux[mm + 1];
            wvy[mm] = wvy[mm + 1];
        }
        wux[7] = ux_;
        wvy[7] = vy_;
    }
    if (inbox) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int64_t o = (int64_t)(kbeg + m) * pg + own;
            wu[m] = q_u[o]; wv[m] = q_v[o]; ww[m] = q_w[o]; wt[m] = q_t[o];
        }
        const int64_t o = (int64_t)(kbeg + 2 * NG) * pg + own;
        nw = q_w[o];
        nt = q_t[o];
    }

    for (int k = kbeg; k < kend; k++) {
        const int cur = (k - kbeg) & 1;
        const double *b = sm + cur * (DNF * S::TILE);
        plane_barrier();
        if (k + 1 < kend)
            dma_plane_s<S, DNF>(src, (int64_t)(k + 1 + NG) * pg, sm + (cur ^ 1) * (DNF * S::TILE), td, wave, lane);
        const double *s_cu = b + DCU * S::TILE, *s_cv = b + DCV * S::TILE, *s_cw = b + DCW * S::TILE,
                     *s_ct = b + DCT * S::TILE, *s_wz = b + DWZ * S::TILE, *s_lu = b + DLU * S::TILE,
                     *s_lv = b + DLV * S::TILE;

        const int64_t zo = (int64_t)(k + NG) * pg + own;
        wu[8] = s_lu[me];
        wv[8] = s_lv[me];
        ww[8] = nw;
        wt[8] = nt;
        if (inbox && k + 1 < kend) {
            const int64_t o = zo + 5 * pg;
            nw = q_w[o];
            nt = q_t[o];
        }

        // ---------------- phase B
        {
#define LUX(mm) s_lu[me + (mm)]
#define LVY(mm) s_lv[me + (mm) * S::PX]
            // leading plane: ux, vy of the own column (cpu/diffterm.cpp:77-81, :110-114)
            wux[8] = D1(LUX) * dxi0;
            wvy[8] = D1(LVY) * dxi1;
#undef LUX
#undef LVY
            // centre plane, ux: the own rows from the window (ux of plane k, computed from
            // the same values four planes ago), the halo rows 0..3 and 12..15 from the tile
            s_ux[me64] = wux[4];
            {
                const int rr = threadIdx.x / S::TX, r = rr < 4 ? rr : rr + S::BY, c = threadIdx.x % S::TX;
                const int l = r * S::PX + c + 4;
#define CUX(mm) s_cu[l + (mm)]
                s_ux[r * S::TX + c] = D1(CUX) * dxi0;
#undef CUX
            }
            // centre plane, vy: the own cells from the window, the x-halo columns 0..3 and
            // 68..71 of rows 4..11 from the tile
            s_vy[meby] = wvy[4];
            if (threadIdx.x < 8 * S::BY) {
                const int rr = threadIdx.x >> 3, cc = threadIdx.x & 7, c = cc < 4 ? cc : cc + S::TX;
                const int l = (rr + 4) * S::PX + c;
#define CVY(mm) s_cv[l + (mm) * S::PX]
                s_vy[rr * S::PX + c] = D1(CVY) * dxi1;
#undef CVY
            }
        }
        __syncthreads();

        // ---------------- phase C
        if (active) {
#define XS(a, mm) a[me + (mm)]
#define YS(a, mm) a[me + (mm) * S::PX]
#define CUx(mm) XS(s_cu, mm)
#define CUy(mm) YS(s_cu, mm)
#define CUz(mm) wu[4 + (mm)]
#define CVx(mm) XS(s_cv, mm)
#define CVy(mm) YS(s_cv, mm)
#define CVz(mm) wv[4 + (mm)]
#define CWx(mm) XS(s_cw, mm)
#define CWy(mm) YS(s_cw, mm)
#define CWz(mm) ww[4 + (mm)]
#define CTx(mm) XS(s_ct, mm)
#define CTy(mm) YS(s_ct, mm)
#define CTz(mm) wt[4 + (mm)]
#define VYx(mm) s_vy[meby + (mm)]
#define WZx(mm) XS(s_wz, mm)
#define UXy(mm) s_ux[me64 + (mm) * S::TX]
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
            if (WRITE_RHO) difflux[fo] = 0.0E0;
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
            wu[m] = wu[m + 1];
            wv[m] = wv[m + 1];
            ww[m] = ww[m + 1];
            wt[m] = wt[m + 1];
            wux[m] = wux[m + 1];
            wvy[m] = wvy[m + 1];
        }
    }
}

// z runs: enough workgroups to fill 256 CUs several times over, with runs long enough
// that the 8-plane window warm-up stays a small overhead; small boxes then shorten the
// runs down to 4 planes until a launch has 256 workgroups (64^3 step 0.60 -> 0.40 ms,
// 32^3 0.55 -> 0.35 ms; at 128^3, with 512 workgroups at runs of 16, shorter runs are
// slower).
// CNS_PIPE_KZ in the environment caps the run length (for measurements).
int pipe_kz(const Geom &g, int tiles_xy, int nbox = 1)
{
    int kz = g.n2;
    static const int cap = [] {
        const char *e = getenv("CNS_PIPE_KZ");
        return e ? atoi(e) : 0;
    }();
    if (cap >= 4 && kz > cap) kz = cap;
    auto blocks = [&](int z) { return (int64_t)tiles_xy * nbox * ((g.n2 + z - 1) / z); };
    while (kz > 16 && blocks(kz) < 2048) kz = (kz + 1) / 2;
    while (kz > 4 && blocks(kz) < 256) kz = (kz + 1) / 2;
    return kz;
}

// The step's stencils run on 32 x 8 tiles, two workgroups per CU (512^3 step: diffterm
// 4.84 -> 4.33 ms, fused hypterm 6.65 -> 6.05 ms against 64 x 8); setting the named
// variable to 64 selects 64 x 8 tiles (the results are the same bits either way).
bool wide_tiles(const char *var)
{
    const char *e = getenv(var);
    return e && atoi(e) == 64;
}

template <class S>
hipError_t launch_hyp_pipe_shape(const Geom &g, const double *cons, const double *q, double *flux, double dxi0,
                                 double dxi1, double dxi2, const double *d, const double *base, double *out,
                                 const double *dt_dev, int phase, bool write_f, hipStream_t s)
{
    const int ntx = (g.n0 + S::TX - 1) / S::TX, nty = (g.n1 + S::BY - 1) / S::BY;
    const int kz = pipe_kz(g, ntx * nty);
    const int nb = ntx * nty * ((g.n2 + kz - 1) / kz);
    PipeRk rk{d, base, out, dt_dev, phase};
    const bool stage1 = phase == 1 && base == cons;
#define HYP_PIPE(RK, S1, WF)                                                                              \
    hypterm_pipe_kernel<S, RK, S1, WF><<<nb, S::NT, 0, s>>>(g, cons, q, flux, dxi0, dxi1, dxi2, kz, ntx, nty, \
                                                            rk)
    if (!phase)
        HYP_PIPE(false, false, true);
    else if (stage1)
        write_f ? HYP_PIPE(true, true, true) : HYP_PIPE(true, true, false);
    else
        write_f ? HYP_PIPE(true, false, true) : HYP_PIPE(true, false, false);
#undef HYP_PIPE
    return hipGetLastError();
}

template <class S>
hipError_t launch_diffterm_shape(const Geom &g, const double *q, const double *wz, double *difflux,
                                 double dxi0, double dxi1, double dxi2, double eta, double alam,
                                 hipStream_t s, bool write_rho, const BoxSet &bs)
{
    const int ntx = (g.n0 + S::TX - 1) / S::TX, nty = (g.n1 + S::BY - 1) / S::BY;
    const int kz = pipe_kz(g, ntx * nty, bs.n);
    const int nbpb = ntx * nty * ((g.n2 + kz - 1) / kz);
    if (bs.n < 1 || (int64_t)nbpb * bs.n > 0x7fffffff) return hipErrorInvalidValue;
    const int nb = nbpb * bs.n;
    if (write_rho)
        diffterm_pipe_kernel<S, true><<<nb, S::NT, 0, s>>>(g, q, wz, difflux, dxi0, dxi1, dxi2, eta, alam, kz,
                                                          ntx, nty, bs, nbpb);
    else
        diffterm_pipe_kernel<S, false><<<nb, S::NT, 0, s>>>(g, q, wz, difflux, dxi0, dxi1, dxi2, eta, alam, kz,
                                                           ntx, nty, bs, nbpb);
    return hipGetLastError();
}

template <class S>
hipError_t launch_hyp_cons_shape(const Geom &g, const double *cons, double *flux, double dxi0, double dxi1,
                                 double dxi2, const double *d, const double *base, double *out,
                                 const double *dt_dev, int phase, bool write_f, double *qout, int qout_kind,
                                 hipStream_t s, const BoxSet &bs)
{
    const int ntx = (g.n0 + S::TX - 1) / S::TX, nty = (g.n1 + S::BY - 1) / S::BY;
    const int kz = pipe_kz(g, ntx * nty, bs.n);
    const int nbpb = ntx * nty * ((g.n2 + kz - 1) / kz);
    if (bs.n < 1 || (int64_t)nbpb * bs.n > 0x7fffffff) return hipErrorInvalidValue;
    const int nb = nbpb * bs.n;
    PipeRk rk{d, base, out, dt_dev, phase, qout};
    const bool stage1 = phase == 1 && base == cons;
    if (!qout) qout_kind = 0;
    if (qout_kind < 0 || qout_kind > 2 || (write_f && qout_kind)) return hipErrorInvalidValue;
#define HYP_CONS(S1, WF, QO) \
    hypterm_cons_kernel<S, S1, WF, QO><<<nb, S::NT, 0, s>>>(g, cons, flux, dxi0, dxi1, dxi2, kz, ntx, nty, rk, bs, nbpb)
    // the step's combinations: stage 1 (-> q 1), stage 2 (-> q 2), stage 3 (F, no q);
    // and no q for any stage
    if (stage1 && !write_f)
        qout_kind == 1 ? HYP_CONS(true, false, 1) : (qout_kind == 2 ? HYP_CONS(true, false, 2) : HYP_CONS(true, false, 0));
    else if (stage1)
        HYP_CONS(true, true, 0);
    else if (!write_f)
        qout_kind == 1 ? HYP_CONS(false, false, 1) : (qout_kind == 2 ? HYP_CONS(false, false, 2) : HYP_CONS(false, false, 0));
    else
        HYP_CONS(false, true, 0);
#undef HYP_CONS
    return hipGetLastError();
}

}  // namespace

bool pipe_aligned(const void *p) { return ((uintptr_t)p & 15) == 0; }

hipError_t launch_hypterm_pipe(const Geom &g, const double *cons, const double *q, double *flux,
                               double dxi0, double dxi1, double dxi2, const double *d,
                               const double *base, double *out, const double *dt_dev, int phase,
                               bool write_f, hipStream_t s)
{
    if (wide_tiles("CNS_HYP_TX"))
        return launch_hyp_pipe_shape<Shape<64>>(g, cons, q, flux, dxi0, dxi1, dxi2, d, base, out, dt_dev, phase,
                                                write_f, s);
    return launch_hyp_pipe_shape<Shape<32>>(g, cons, q, flux, dxi0, dxi1, dxi2, d, base, out, dt_dev, phase,
                                            write_f, s);
}

hipError_t launch_hypterm_cons(const Geom &g, const double *cons, double *flux, double dxi0,
                               double dxi1, double dxi2, const double *d, const double *base,
                               double *out, const double *dt_dev, int phase, bool write_f,
                               double *qout, int qout_kind, hipStream_t s, const BoxSet &bs)
{
    if (wide_tiles("CNS_HYP_TX"))
        return launch_hyp_cons_shape<Shape<64>>(g, cons, flux, dxi0, dxi1, dxi2, d, base, out, dt_dev, phase,
                                                write_f, qout, qout_kind, s, bs);
    return launch_hyp_cons_shape<Shape<32>>(g, cons, flux, dxi0, dxi1, dxi2, d, base, out, dt_dev, phase,
                                            write_f, qout, qout_kind, s, bs);
}


// z runs of the column kernels (ctoprim + wz, wz): long runs keep the 8-plane window
// warm-up small; small boxes shorten them, down to 8 planes (or the starting run when
// that is shorter), until a launch has 512 workgroups (at 64^3 the ctoprim + wz pass
// had 46 workgroups of 64-plane runs).
static int column_kz(int kz, int64_t col_blocks, int nz, int nbox)
{
    while (kz > 8 && col_blocks * nbox * ((nz + kz - 1) / kz) < 512) kz = kz / 2 > 8 ? kz / 2 : 8;
    return kz;
}

hipError_t launch_ctoprim_wz(const Geom &g, const double *u, double *q, double *wz,
                             unsigned long long *acc, double dx0, double dx1, double dx2, double dxi2,
                             bool full, hipStream_t s, const BoxSet &bs)
{
    if (bs.n < 1 || bs.n > 65535) return hipErrorInvalidValue;
    const int64_t cols = (g.pg + kCtoBlock - 1) / kCtoBlock;
    const int kz = column_kz(64, cols, g.g2, bs.n);
    const dim3 grid((unsigned)cols, (unsigned)((g.g2 + kz - 1) / kz), (unsigned)bs.n);
#define CTO_WZ(C, F) \
    ctoprim_wz_kernel<C, F><<<grid, kCtoBlock, 0, s>>>(g, u, q, wz, acc, dx0, dx1, dx2, dxi2, kz, bs)
    if (acc)
        full ? CTO_WZ(true, true) : CTO_WZ(true, false);
    else
        full ? CTO_WZ(false, true) : CTO_WZ(false, false);
#undef CTO_WZ
    return hipGetLastError();
}

hipError_t launch_wz(const Geom &g, const double *q, double *wz, double dxi2, hipStream_t s, const BoxSet &bs)
{
    if (bs.n < 1 || bs.n > 65535) return hipErrorInvalidValue;
    const int64_t cols = (g.pg + 255) / 256;
    const int kz = column_kz(g.n2 > 128 ? 128 : g.n2, cols, g.n2, bs.n);
    const dim3 grid((unsigned)cols, (unsigned)((g.n2 + kz - 1) / kz), (unsigned)bs.n);
    wz_kernel<<<grid, 256, 0, s>>>(g, q + QW * g.cg, wz, dxi2, kz, bs);
    return hipGetLastError();
}

hipError_t launch_diffterm_pipe(const Geom &g, const double *q, const double *wz, double *difflux,
                                double dxi0, double dxi1, double dxi2, double eta, double alam,
                                hipStream_t s, bool write_rho, const BoxSet &bs)
{
    if (wide_tiles("CNS_DIFF_TX"))
        return launch_diffterm_shape<Shape<64>>(g, q, wz, difflux, dxi0, dxi1, dxi2, eta, alam, s, write_rho, bs);
    return launch_diffterm_shape<Shape<32>>(g, q, wz, difflux, dxi0, dxi1, dxi2, eta, alam, s, write_rho, bs);
}

}  // namespace cns
