// Host build of kf_split (killingfusion_amd/csrc/kf_internal.hpp): reads "i r" pairs (float r
// as its bit pattern in hex) from stdin and prints "b f" (f as its bit pattern in hex), so
// tests/test_split_cpu.py can check the split against the reference's read convention.
#include "kf_internal.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>

int main()
{
    int i;
    unsigned int rb;
    while (std::scanf("%d %x", &i, &rb) == 2) {
        float r;
        std::memcpy(&r, &rb, 4);
        int b;
        float f;
        kf_split(i, r, b, f);
        unsigned int fb;
        std::memcpy(&fb, &f, 4);
        std::printf("%d %08x\n", b, fb);
    }
    return 0;
}
