// Host stand-in for the HIP runtime header, used only to compile the
// double-float routines of covertreex_amd/csrc/df.hpp on the CPU for their
// accuracy test: the device qualifiers become no-ops and the two hardware
// approximations (v_rcp_f32, v_sqrt_f32, about 1 ulp) are modelled as the
// correctly rounded value moved by one ulp in a direction chosen per call.
#pragma once
#include <cmath>
#include <cstdint>
#define __device__
#define __forceinline__ inline
#define __constant__
extern uint32_t df_host_ulp_state;
inline float df_host_nudge(float v)
{
    df_host_ulp_state = df_host_ulp_state * 1664525u + 1013904223u;
    const uint32_t r = df_host_ulp_state >> 30;
    return r == 0 ? std::nextafter(v, INFINITY) : r == 1 ? std::nextafter(v, -INFINITY) : v;
}
inline float __builtin_amdgcn_rcpf(float x) { return df_host_nudge(1.0f / x); }
inline float __builtin_amdgcn_sqrtf(float x) { return df_host_nudge(std::sqrt(x)); }
