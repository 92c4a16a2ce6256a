// Correctly rounded binary32 exp(x) in FP32 arithmetic (no FP64), for the
// build and the reference-form metric.
//
// The reference evaluates its kernel values with the platform expf
// (src/metric.rs:183,196,458,464,534,541); the tree and the reference-form
// query here use the correctly rounded exp, whose trees equal the libm ones
// on the workloads checked.  Over every float x in [-110, 89] exp(x) lies at
// least 2.36e-9 ulp from a rounding midpoint (the closest is
// x = -0x1.d2259ap+3), and this routine returns the correctly rounded value
// for each of them (checked exhaustively on the CPU by the generator and on
// the GPU against the oracle).
//
// Method (Tang's table-driven exp, with a Ziv-style exact rounding test):
//   x = (256 k + j) ln2/256 + r, N = 256 k + j = rint(x 256/ln2), |r| <= 2^-9.4,
//   ln2/256 as four floats (the first with 8 significant bits, so N C1 is
//   exact for |N| < 2^16), r kept as a double-float;
//   exp(x) = 2^k 2^(j/256) exp(r), 2^(j/256) as three floats (Th + Tm + Tl,
//   72 bits, generated by gen_exp_cr.cpp), and
//   y = Th + [Th r + Th r^2/2 + Tm] + small, with the products Th rh and
//   Th rh^2/2 exact (fma TwoProduct) and the small terms (r^3/6 + r^4/24 +
//   r^5/120, the low parts, Tm r, Tl; below 2^-30) in FP32.  y is held as the
//   unevaluated sum s + t + c with s = RN(y), t the exact TwoSum remainder and
//   c the small remainder; whether y lies above or below the rounding
//   midpoint next to s is decided exactly (Sterbenz differences and the sign
//   of one rounded sum).  Results below 2^-126 are rounded on the subnormal
//   grid from the integer significand of s and the sign of t + c.
//   The approximation error of s + t + c is below kExpCrErr = 2^-28 result
//   ulps over the whole range (2.7e-9 at most, measured exhaustively by the
//   generator); an input whose y lies within kExpCrGate ulps of a midpoint
//   (5 of the 2.24e9) is looked up in the list of the 18 inputs whose exact
//   exp lies within kExpCrHard ulps of one (binary search); for every other
//   input the rounding decision is certain.
#pragma once

#include "exp_cr_core.hpp"
#include "exp_cr_table.inc"

namespace cx {

// exp(x) correctly rounded to binary32 (x > 89: +inf, x < -110: +0, NaN
// passes through)
__device__ __forceinline__ float expf_cr(float x)
{
    if (x != x)
        return x;
    if (x > 89.0f)
        return __builtin_huge_valf();
    if (x < -110.0f)
        return 0.0f;
    bool gate;
    const float r = ecr::round_core(ecr::core(x, kExpCrTab), kExpCrGate, &gate);
    if (gate) {
        // the listed inputs whose exp lies within kExpCrHard ulps of a
        // midpoint, sorted by bit pattern
        const uint32_t xb = ecr::bits(x);
        int lo = 0, hi = kExpCrHardN;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (kExpCrHardX[mid] < xb)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < kExpCrHardN && kExpCrHardX[lo] == xb)
            return ecr::from_bits(kExpCrHardY[lo]);
    }
    return r;
}

} // namespace cx
