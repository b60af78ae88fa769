// Host build of U32Div (killingfusion_amd/csrc/kf_internal.hpp), the multiply-high division by
// the grid's row and slice sizes that every voxel-index decode uses: reads "d n" pairs from
// stdin and prints "q r" (n / d and n - q d as U32Div computes them), so that
// tests/test_udiv_cpu.py can compare them with exact integer division.
#include "kf_internal.hpp"

#include <cstdio>

int main()
{
    unsigned int d, n;
    while (std::scanf("%u %u", &d, &n) == 2) {
        const U32Div v = U32Div::of(d);
        const unsigned int q = v.div(n);
        std::printf("%u %u\n", q, n - q * d);
    }
    return 0;
}
