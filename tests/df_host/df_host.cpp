// Accuracy of the double-float routines of df.hpp on the CPU: prints, per
// routine, the largest relative error against long double over random
// arguments (IEEE float arithmetic, no contraction: built with
// -ffp-contract=off, as the device code).
#include <cstdio>
#include <random>
#include "df.hpp"

uint32_t df_host_ulp_state = 12345u;
using namespace cx;

static long double val(df a) { return (long double)a.hi + (long double)a.lo; }

static df make_df(std::mt19937_64 &g, double lo, double hi)
{
    std::uniform_real_distribution<double> u(lo, hi);
    const double v = u(g);
    const float h = (float)v;
    return df{h, (float)(v - (double)h)};
}

int main()
{
    std::mt19937_64 g(7);
    long double e_exp = 0, e_exp_abs = 0, e_div = 0, e_sqrt = 0, e_add = 0, e_mul = 0;
    for (int i = 0; i < 400000; i++) {
        const df x = make_df(g, -103.0, 0.0);
        const long double ex = expl(val(x));
        const long double r = val(df_exp(x));
        // relative while the low part is a normal float; absolute below
        if (ex > 1e-30L)
            e_exp = std::max(e_exp, fabsl(r - ex) / ex);
        else
            e_exp_abs = std::max(e_exp_abs, fabsl(r - ex));
        const df a = make_df(g, -4.0, 4.0), b = make_df(g, 1e-3, 8.0);
        const long double q = val(a) / val(b);
        if (q != 0)
            e_div = std::max(e_div, fabsl(val(df_div(a, b)) - q) / fabsl(q));
        const df c = make_df(g, 1e-12, 2.0);
        const long double sq = sqrtl(val(c));
        e_sqrt = std::max(e_sqrt, fabsl(val(df_sqrt(c)) - sq) / sq);
        const long double sm = val(a) + val(b);
        e_add = std::max(e_add, fabsl(val(df_add(a, b)) - sm) / (fabsl(val(a)) + fabsl(val(b))));
        const long double pr = val(a) * val(b);
        if (pr != 0)
            e_mul = std::max(e_mul, fabsl(val(df_mul(a, b)) - pr) / fabsl(pr));
    }
    printf("exp %.6Le\nexp_abs %.6Le\ndiv %.6Le\nsqrt %.6Le\nadd %.6Le\nmul %.6Le\n", e_exp,
           e_exp_abs, e_div, e_sqrt, e_add, e_mul);
    return 0;
}
