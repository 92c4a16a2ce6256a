// expf_cr (covertreex_amd/csrc/exp_cr.hpp) compiled for the CPU: reads
// binary32 bit patterns (uint32, native order) from argv[1] and writes
// expf_cr of each as float32 to argv[2].  Built with -ffp-contract=off, as the
// device code.
#include <cstdint>
#include <cstdio>
#include <vector>

#define __device__
#define __forceinline__ inline
#define __constant__
#include "exp_cr.hpp"

int main(int argc, char **argv)
{
    if (argc != 3)
        return 2;
    FILE *in = std::fopen(argv[1], "rb");
    if (!in)
        return 1;
    std::vector<uint32_t> x;
    uint32_t b;
    while (std::fread(&b, 4, 1, in) == 1)
        x.push_back(b);
    std::fclose(in);
    std::vector<float> y(x.size());
    for (size_t i = 0; i < x.size(); i++)
        y[i] = cx::expf_cr(__builtin_bit_cast(float, x[i]));
    FILE *out = std::fopen(argv[2], "wb");
    if (!out)
        return 1;
    std::fwrite(y.data(), 4, y.size(), out);
    std::fclose(out);
    return 0;
}
