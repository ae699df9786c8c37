// cns_stencil_ops.h -- 8th-order difference operators shared by the stencil kernels.
//
// Each macro takes an offset accessor F(m) (m = -4..4 along one direction) and
// expands to the reference's expression for that term, keeping its operation
// order (cpu/hypterm.cpp:43-90, cpu/diffterm.cpp:77-81 and :159-163).  KK is a
// constant set of the arithmetic type in use: K<double> reproduces the
// reference's constants exactly; K<float> rounds them once to float.
#pragma once

#include "cns_common.h"

namespace cns {

template <typename T>
struct K {
    static constexpr T ALP = T(0.8E0);
    static constexpr T BET = T(-0.2E0);
    static constexpr T GAM = T(4.0E0 / 105.0E0);
    static constexpr T DEL = T(-1.0E0 / 280.0E0);
    static constexpr T CENTER = T(-205.0E0 / 72.0E0);
    static constexpr T OFF1 = T(8.0E0 / 5.0E0);
    static constexpr T OFF2 = T(-0.2E0);
    static constexpr T OFF3 = T(8.0E0 / 315.0E0);
    static constexpr T OFF4 = T(-1.0E0 / 560.0E0);
    static constexpr T OneThird = T(1.0E0 / 3.0E0);
    static constexpr T TwoThirds = T(2.0E0 / 3.0E0);
    static constexpr T FourThirds = T(4.0E0 / 3.0E0);
    static constexpr T Two = T(2.0E0);
};

// acc + a*b.  In the flux type float it is one fused multiply-add (one rounding, one
// VALU instruction); in double it is the reference's separate product and sum, so the
// FP64 path keeps the reference's bits (the library is compiled without contraction).
__device__ __forceinline__ float mac(float acc, float a, float b) { return __builtin_fmaf(a, b, acc); }
__device__ __forceinline__ double mac(double acc, double a, double b) { return acc + a * b; }

}  // namespace cns

// The operators below are chains of mac() in the reference's left-to-right order.  In
// double every step rounds exactly as the reference's expression does (x - y == -y + x
// in IEEE arithmetic), so the FP64 results are bit-identical to the unfused forms.

// First difference sum (before *dxinv).
#define KD1(F)                                                                                 \
    cns::mac(cns::mac(cns::mac(KK::ALP * (F(1) - F(-1)), KK::BET, F(2) - F(-2)), KK::GAM,     \
                      F(3) - F(-3)),                                                          \
             KK::DEL, F(4) - F(-4))
// Second difference sum (before *dxinv^2).
#define KD2(F)                                                                                 \
    cns::mac(cns::mac(cns::mac(cns::mac(KK::CENTER * F(0), KK::OFF1, F(1) + F(-1)), KK::OFF2, \
                               F(2) + F(-2)),                                                 \
                      KK::OFF3, F(3) + F(-3)),                                                \
             KK::OFF4, F(4) + F(-4))
// C(m)*UN(m) - C(-m)*UN(-m)
#define KPD(C, UN, m) cns::mac(-(C(-(m)) * UN(-(m))), C(m), UN(m))
// hypterm: tangential momentum flux sum, normal momentum (with pressure), energy.
#define KHSUM(T)                                                                               \
    cns::mac(cns::mac(cns::mac(KK::ALP * (T(1)), KK::BET, T(2)), KK::GAM, T(3)), KK::DEL, T(4))
#define KHMOM(C, UN) KHSUM([&](int m_) { return KPD(C, UN, m_); })
#define KHMOMP(C, UN, P) KHSUM([&](int m_) { return KPD(C, UN, m_) + (P(m_) - P(-m_)); })
#define KHENE(C, UN, P) KHSUM([&](int m_) { return KPD(C, UN, m_) + KPD(P, UN, m_); })
