// Skilling Hilbert code of one point (covertreex/algo/order/strategy.py:38-99):
// per-axis quantisation to `bits` bits with round-half-even in FP64, the
// AxesToTranspose transform, and the bit interleave, most significant first.
// Shared by the batch order and the exact Euclidean kNN's seed search.
#pragma once

#include <stdint.h>

namespace cx {

// round((x - lo) / max(hi - lo, 1e-9) * (2^bits - 1)) clipped to [0, 2^bits - 1]
__device__ __forceinline__ uint64_t hilbert_quantize(double x, double lo, double hi,
                                                     int32_t bits)
{
    const uint64_t max_val = (1ull << bits) - 1;
    double span = hi - lo;
    span = span > 1e-9 ? span : 1e-9; // np.maximum(hi - lo, 1e-9)
    double r = __builtin_rint((x - lo) / span * (double)max_val); // np.round: half to even
    r = r < 0.0 ? 0.0 : r;
    r = r > (double)max_val ? (double)max_val : r;
    return (uint64_t)r;
}

// v[0..d) are the quantised coordinates (overwritten)
__device__ __forceinline__ uint64_t hilbert_from_quantized(uint64_t *v, int32_t d, int32_t bits)
{
    // Skilling AxesToTranspose (strategy.py:63-88)
    const uint64_t M = 1ull << (bits - 1);
    for (uint64_t Q = M; Q > 1; Q >>= 1) {
        const uint64_t P = Q - 1;
        for (int32_t c = 0; c < d; c++) {
            if (v[c] & Q) {
                v[0] ^= P;
            } else {
                uint64_t t = (v[0] ^ v[c]) & P;
                v[0] ^= t;
                v[c] ^= t;
            }
        }
    }
    for (int32_t c = 1; c < d; c++)
        v[c] ^= v[c - 1];
    uint64_t t = 0;
    for (uint64_t Q = M; Q > 1; Q >>= 1)
        if (v[d - 1] & Q)
            t ^= Q - 1;
    for (int32_t c = 0; c < d; c++)
        v[c] ^= t;
    // interleave, most significant bit first (strategy.py:91-99)
    uint64_t code = 0;
    for (int32_t b = bits - 1; b >= 0; b--) {
        uint64_t word = 0;
        for (int32_t c = 0; c < d; c++)
            word = (word << 1) | ((v[c] >> b) & 1ull);
        code = (code << d) | word;
    }
    return code;
}

__host__ __device__ inline int32_t hilbert_bits(int32_t d)
{
    int32_t bits = 64 / d;
    return bits > 16 ? 16 : (bits < 1 ? 1 : bits);
}

} // namespace cx
