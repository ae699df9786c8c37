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

}  // namespace cns

// First difference sum (before *dxinv).
#define KD1(F)                                                                      \
    (KK::ALP * (F(1) - F(-1)) + KK::BET * (F(2) - F(-2)) + KK::GAM * (F(3) - F(-3)) \
     + KK::DEL * (F(4) - F(-4)))
// Second difference sum (before *dxinv^2).
#define KD2(F)                                                                              \
    (KK::CENTER * F(0) + KK::OFF1 * (F(1) + F(-1)) + KK::OFF2 * (F(2) + F(-2))            \
     + KK::OFF3 * (F(3) + F(-3)) + KK::OFF4 * (F(4) + F(-4)))
// hypterm: tangential momentum flux sum, normal momentum (with pressure), energy.
#define KHMOM(C, UN)                                                                 \
    (KK::ALP * (C(1) * UN(1) - C(-1) * UN(-1)) + KK::BET * (C(2) * UN(2) - C(-2) * UN(-2)) \
     + KK::GAM * (C(3) * UN(3) - C(-3) * UN(-3)) + KK::DEL * (C(4) * UN(4) - C(-4) * UN(-4)))
#define KHMOMP(C, UN, P)                                                              \
    (KK::ALP * (C(1) * UN(1) - C(-1) * UN(-1) + (P(1) - P(-1)))                       \
     + KK::BET * (C(2) * UN(2) - C(-2) * UN(-2) + (P(2) - P(-2)))                     \
     + KK::GAM * (C(3) * UN(3) - C(-3) * UN(-3) + (P(3) - P(-3)))                     \
     + KK::DEL * (C(4) * UN(4) - C(-4) * UN(-4) + (P(4) - P(-4))))
#define KHENE(C, UN, P)                                                               \
    (KK::ALP * (C(1) * UN(1) - C(-1) * UN(-1) + (P(1) * UN(1) - P(-1) * UN(-1)))      \
     + KK::BET * (C(2) * UN(2) - C(-2) * UN(-2) + (P(2) * UN(2) - P(-2) * UN(-2)))    \
     + KK::GAM * (C(3) * UN(3) - C(-3) * UN(-3) + (P(3) * UN(3) - P(-3) * UN(-3)))    \
     + KK::DEL * (C(4) * UN(4) - C(-4) * UN(-4) + (P(4) * UN(4) - P(-4) * UN(-4))))
