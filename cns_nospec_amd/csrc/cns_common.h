// cns_common.h -- shared definitions of the CNS_NoSpec HIP library (gfx950).
//
// Layout of every array the C ABI takes (gpu/header.h:14, gpu/util_cpu.cpp:327-343):
//   ghosted  (U, Unew: 5 comps, Q: 6 comps): idx = l*cg + k*pg + j*xg + i,
//            xg = PAD16(n0+8), pg = xg*(n1+8), cg = pg*(n2+8), 0 <= i,j,k < n+8
//   interior (D, F: 5 comps):                idx = l*c + k*p + j*x + i,
//            x = PAD16(n0), p = x*n1, c = p*n2, 0 <= i,j,k < n
// Padding lanes are never read or written.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cns {

constexpr int NG = 4;  // ghost cells (cpu/header.h:16)
constexpr int NC = 5;  // conserved components
constexpr int NQ = 6;  // primitives
enum { IRHO = 0, IMX, IMY, IMZ, IENE };
enum { QRHO = 0, QU, QV, QW, QPRES, QTEMP };

// Stencil weights (cpu/header.h:88-91, cpu/diffterm.cpp:48-52) and ideal-gas
// constants (cpu/ctoprim.cpp:23-24).  All are folded at compile time with
// correctly rounded double arithmetic, the same values the reference computes.
constexpr double ALP = 0.8E0;
constexpr double BET = -0.2E0;
constexpr double GAM = 4.0E0 / 105.0E0;
constexpr double DEL = -1.0E0 / 280.0E0;
constexpr double CENTER = -205.0E0 / 72.0E0;
constexpr double OFF1 = 8.0E0 / 5.0E0;
constexpr double OFF2 = -0.2E0;
constexpr double OFF3 = 8.0E0 / 315.0E0;
constexpr double OFF4 = -1.0E0 / 560.0E0;
constexpr double OneThird = 1.0E0 / 3.0E0;
constexpr double TwoThirds = 2.0E0 / 3.0E0;
constexpr double FourThirds = 4.0E0 / 3.0E0;
constexpr double OneQuarter = 1.E0 / 4.E0;
constexpr double ThreeQuarters = 3.E0 / 4.E0;
constexpr double GAMMA = 1.4E0;
constexpr double CV = 8.3333333333E6;

inline int pad16(int x) { return (((x - 1) >> 4) + 1) << 4; }

// Geometry of one box, passed to kernels by value.
struct Geom {
    int n0, n1, n2;  // interior extents
    int g0, g1, g2;  // ghosted extents
    int xg;          // ghosted x pitch
    int x;           // interior x pitch
    int64_t pg, cg;  // ghosted plane / component pitch
    int64_t p, c;    // interior plane / component pitch
};

inline Geom make_geom(int n0, int n1, int n2)
{
    Geom g;
    g.n0 = n0; g.n1 = n1; g.n2 = n2;
    g.g0 = n0 + 2 * NG; g.g1 = n1 + 2 * NG; g.g2 = n2 + 2 * NG;
    g.xg = pad16(g.g0);
    g.pg = (int64_t)g.xg * g.g1;
    g.cg = g.pg * g.g2;
    g.x = pad16(n0);
    g.p = (int64_t)g.x * n1;
    g.c = g.p * n2;
    return g;
}

// Ghosted offset of cell (i,j,k) in ghost-shifted indices, and interior offset of
// interior cell (i,j,k) counted from 0.
__device__ __forceinline__ int64_t gidx(const Geom &g, int i, int j, int k)
{
    return (int64_t)k * g.pg + (int64_t)j * g.xg + i;
}
__device__ __forceinline__ int64_t iidx(const Geom &g, int i, int j, int k)
{
    return (int64_t)k * g.p + (int64_t)j * g.x + i;
}

// Courant maxima are kept as the bit pattern of a non-negative double in a
// uint64 slot: for such values integer order equals floating order, so an
// integer atomicMax is an exact floating max.  Any NaN is stored as the
// canonical quiet NaN 0x7ff8..., which orders above +inf and so survives the
// max; the host turns it into an error.
// ctoprim of one cell (cpu/ctoprim.cpp:80-100): the single definition of the
// primitive arithmetic, shared by every kernel that derives primitives, so that all
// of them produce the same bits.
__device__ __forceinline__ void ctoprim_cell(double rho, double mx, double my, double mz, double E,
                                             double &u, double &v, double &w, double &eint, double &p)
{
    const double rhoinv = 1.0E0 / rho;
    u = mx * rhoinv;
    v = my * rhoinv;
    w = mz * rhoinv;
    eint = E * rhoinv - 0.5E0 * ((u * u) + (v * v) + (w * w));
    p = (GAMMA - 1.0E0) * eint * rho;
}

__device__ __forceinline__ unsigned long long courno_bits(double v)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (!(v == v)) b = 0x7ff8000000000000ull;
    return b;
}

}  // namespace cns
