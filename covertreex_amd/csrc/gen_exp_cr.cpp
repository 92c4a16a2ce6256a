// Generates exp_cr_table.inc, the tables of the correctly rounded binary32
// exp (exp_cr.hpp), and checks the routine over every float in [-110, 89].
//
//   g++ -O2 -std=c++17 -ffp-contract=off -march=native -pthread \
//       gen_exp_cr.cpp -lquadmath -o gen_exp_cr && ./gen_exp_cr exp_cr_table.inc
//
// 1. 2^(j/256), j < 256, in binary128 (exp2q), split into three floats.
// 2. For every float x in [-110, 89] (2.24e9 inputs, all host threads):
//    exp(x) in x87 extended precision (expl, 64-bit significand: within
//    2^-63 relative, about 2^-40 result ulps, far below the margins here),
//    the routine's unevaluated sum s + t + c and its error in result ulps,
//    and the distance of exp(x) from the nearest rounding midpoint (at least
//    2.36e-9 ulps, so rounding expl(x) gives the correctly rounded value).
//    The error bound kExpCrErr is the largest error
//    rounded up to a power of two; the gate is 2 kExpCrErr and the listed
//    hard inputs are those within 4 kExpCrErr of a midpoint, with their
//    correctly rounded values.
// 3. The routine with these tables (fast path, gate, list) against the
//    correctly rounded value on every input: the program fails unless they
//    all agree.
#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#define __device__
#define __forceinline__ inline
#define __constant__
#include "exp_cr_core.hpp"

using namespace cx;

static float f_of(uint32_t b)
{
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

static uint32_t b_of(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// exponent of the binary32 grid a positive value rounds on
static int grid_exp(long double v)
{
    int e;
    frexpl(v, &e); // v = f 2^e, f in [0.5, 1)
    return std::max(e - 24, -149);
}

// binary32 rounding of a positive value (nearest, ties to even), subnormals
// included
static float round_f32(long double v)
{
    if (v <= 0)
        return 0.0f;
    const int q = grid_exp(v);
    const long double scaled = ldexpl(v, -q); // integer part = significand
    long double ip = floorl(scaled);
    const long double fr = scaled - ip;
    if (fr > 0.5L || (fr == 0.5L && fmodl(ip, 2.0L) != 0))
        ip += 1;
    const long double r = ldexpl(ip, q);
    if (r > 3.4028234663852886e38L)
        return INFINITY;
    return (float)r; // exact: r is a binary32 value
}

// distance of v from the nearest binary32 rounding midpoint, in ulps of the
// grid v rounds on
static double midpoint_distance(long double v)
{
    const long double scaled = ldexpl(v, -grid_exp(v));
    return (double)fabsl(scaled - floorl(scaled) - 0.5L);
}

struct Hard {
    uint32_t x;
    double dist;
    uint32_t y;
};

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s exp_cr_table.inc\n", argv[0]);
        return 2;
    }
    static ExpCrTab tab;
    for (int j = 0; j < 256; j++) {
        const __float128 v = exp2q((__float128)j / 256);
        const float h = (float)(double)v;
        const float m = (float)(double)(v - (__float128)h);
        const float l = (float)(double)(v - (__float128)h - (__float128)m);
        tab.h[j] = h;
        tab.m[j] = m;
        tab.l[j] = l;
    }

    const uint32_t neg_end = b_of(-110.0f), pos_end = b_of(89.0f);
    const uint64_t n_neg = (uint64_t)neg_end - 0x80000000u + 1, n_pos = (uint64_t)pos_end + 1;
    const uint64_t total = n_neg + n_pos;
    auto input = [&](uint64_t i) -> uint32_t {
        return i < n_neg ? (uint32_t)(0x80000000u + i) : (uint32_t)(i - n_neg);
    };
    const unsigned n_thr = std::max(1u, std::thread::hardware_concurrency());

    // pass 1: largest error of s + t + c; inputs within 2^-20 ulp of a midpoint
    const double kCollect = std::ldexp(1.0, -20);
    std::vector<double> err_max(n_thr, 0.0);
    std::vector<std::vector<Hard>> near(n_thr);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < n_thr; t++)
            th.emplace_back([&, t] {
                for (uint64_t i = t; i < total; i += n_thr) {
                    const float x = f_of(input(i));
                    const long double ex = expl((long double)x);
                    if (ex > 3.4028234663852886e38L)
                        continue; // overflows either way
                    const ecr::Core c = ecr::core(x, tab);
                    // s + t + c in extended precision (within 2^-64 relative)
                    const long double approx =
                        ldexpl((long double)c.s + (long double)c.t + (long double)c.c, c.k);
                    const double err = (double)fabsl(ldexpl(approx - ex, -grid_exp(ex)));
                    err_max[t] = std::max(err_max[t], err);
                    const double d = midpoint_distance(ex);
                    if (d < kCollect)
                        near[t].push_back({b_of(x), d, b_of(round_f32(ex))});
                }
            });
        for (auto &x : th)
            x.join();
    }
    double emax = 0.0;
    for (double e : err_max)
        emax = std::max(emax, e);
    int p = -40;
    while (std::ldexp(1.0, p) < emax)
        p++;
    const float err_bound = (float)std::ldexp(1.0, p);
    const float gate = 2.0f * err_bound, hard_width = 4.0f * err_bound;
    std::vector<Hard> hard;
    for (auto &v : near)
        for (const Hard &h : v)
            if (h.dist < hard_width)
                hard.push_back(h);
    std::sort(hard.begin(), hard.end(), [](const Hard &a, const Hard &b) { return a.x < b.x; });
    std::printf("inputs %llu, largest error %.3e ulp (bound 2^%d), hard inputs %zu\n",
                (unsigned long long)total, emax, p, hard.size());
    if (hard_width > kCollect) {
        std::fprintf(stderr, "error bound too large\n");
        return 1;
    }

    // pass 2: the routine (fast path, gate, list) against the correctly
    // rounded value on every input
    std::vector<uint64_t> bad(n_thr, 0), gated(n_thr, 0);
    std::mutex mu;
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < n_thr; t++)
            th.emplace_back([&, t] {
                for (uint64_t i = t; i < total; i += n_thr) {
                    const float x = f_of(input(i));
                    bool g;
                    float r = ecr::round_core(ecr::core(x, tab), gate, &g);
                    if (g) {
                        gated[t]++;
                        const uint32_t xb = b_of(x);
                        auto it = std::lower_bound(
                            hard.begin(), hard.end(), xb,
                            [](const Hard &h, uint32_t v) { return h.x < v; });
                        if (it != hard.end() && it->x == xb)
                            r = f_of(it->y);
                    }
                    const float want = round_f32(expl((long double)x));
                    if (b_of(r) != b_of(want)) {
                        bad[t]++;
                        std::lock_guard<std::mutex> lk(mu);
                        if (bad[t] < 5)
                            std::printf("mismatch x=%a got %a want %a\n", x, r, want);
                    }
                }
            });
        for (auto &x : th)
            x.join();
    }
    uint64_t n_bad = 0, n_gated = 0;
    for (unsigned t = 0; t < n_thr; t++) {
        n_bad += bad[t];
        n_gated += gated[t];
    }
    std::printf("gated %llu, mismatches %llu\n", (unsigned long long)n_gated,
                (unsigned long long)n_bad);
    if (n_bad)
        return 1;

    FILE *f = std::fopen(argv[1], "w");
    if (!f)
        return 1;
    std::fprintf(f, "// Generated by gen_exp_cr.cpp: tables of expf_cr (exp_cr.hpp).\n");
    std::fprintf(f, "// 2^(j/256) = h + m + l (binary128 values split into three floats);\n");
    std::fprintf(f, "// largest error of the unevaluated sum over [-110, 89]: %.3e ulp.\n", emax);
    std::fprintf(f, "namespace cx {\n\n");
    std::fprintf(f, "constexpr float kExpCrErr = %a;  // 2^%d result ulps\n", err_bound, p);
    std::fprintf(f, "constexpr float kExpCrGate = %a; // 2 kExpCrErr\n", gate);
    std::fprintf(f, "constexpr float kExpCrHard = %a; // 4 kExpCrErr\n\n", hard_width);
    std::fprintf(f, "__constant__ constexpr ExpCrTab kExpCrTab = {\n");
    const float *parts[3] = {tab.h, tab.m, tab.l};
    for (int q = 0; q < 3; q++) {
        std::fprintf(f, "    {");
        for (int j = 0; j < 256; j++)
            std::fprintf(f, "%s%a", j ? (j % 4 ? ", " : ",\n     ") : "", parts[q][j]);
        std::fprintf(f, "},\n");
    }
    std::fprintf(f, "};\n\n");
    const size_t nh = hard.size();
    std::fprintf(f, "// inputs (bit patterns, ascending) whose exp lies within kExpCrHard\n");
    std::fprintf(f, "// ulps of a rounding midpoint, and their correctly rounded values\n");
    std::fprintf(f, "constexpr int kExpCrHardN = %zu;\n", nh);
    std::fprintf(f, "__constant__ constexpr uint32_t kExpCrHardX[%zu] = {", std::max<size_t>(nh, 1));
    for (size_t i = 0; i < nh; i++)
        std::fprintf(f, "%s0x%08xu", i ? (i % 6 ? ", " : ",\n    ") : "\n    ", hard[i].x);
    std::fprintf(f, "%s};\n", nh ? "\n" : "0u");
    std::fprintf(f, "__constant__ constexpr uint32_t kExpCrHardY[%zu] = {", std::max<size_t>(nh, 1));
    for (size_t i = 0; i < nh; i++)
        std::fprintf(f, "%s0x%08xu", i ? (i % 6 ? ", " : ",\n    ") : "\n    ", hard[i].y);
    std::fprintf(f, "%s};\n\n} // namespace cx\n", nh ? "\n" : "0u");
    std::fclose(f);
    return 0;
}
