// Host check of the LDS exchange position maps of cc_device.hpp (xswz): for
// every FFT shape the kernels instantiate and every exchange of it, the map
// of that exchange is a bijection of [0, N), is linear over GF(2) (the
// kernels form positions as xswz(a) ^ xswz(b)), and, for 16-byte elements,
// the 8 consecutive lanes of a Stockham write hit 8 distinct bank slots
// (position mod 8: 8 lanes x 16 bytes over 32 banks of 4 bytes per LDS
// cycle).  Prints one line per exchange and exits non-zero on a failure.
#include "../../parallel-critical-collapse_amd/csrc/cc_device.hpp"

#include <cstdio>
#include <set>

// the kernels' own map (cc_device.hpp), compiled for the host
template <int N, typename R, int NS, int RAD>
int host_xswz(int i)
{
    return ccd::xswz<N, R, NS, RAD>(i);
}

template <int N, int T, typename R, int NS>
int check()
{
    constexpr int P = N / T, RAD = (N / NS >= P) ? P : N / NS;
    if constexpr (NS * RAD == N) {
        return 0;
    } else {
        int bad = 0;
        std::set<int> img;
        for (int i = 0; i < N; ++i) img.insert(host_xswz<N, R, NS, RAD>(i));
        if ((int)img.size() != N || *img.begin() != 0 || *img.rbegin() != N - 1) ++bad;
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; b += 5)
                if (host_xswz<N, R, NS, RAD>(a ^ b) != (host_xswz<N, R, NS, RAD>(a) ^ host_xswz<N, R, NS, RAD>(b))) {
                    ++bad;
                    a = N;
                    break;
                }
        int shared = 0;
        if (sizeof(R) == 8)
            for (int g = 0; g < T; g += 8)
                for (int r = 0; r < RAD; ++r) {
                    std::set<int> slots;
                    for (int l = g; l < g + 8; ++l) {
                        const int k = l & (NS - 1);
                        slots.insert((host_xswz<N, R, NS, RAD>((l / NS) * NS * RAD + k) ^ host_xswz<N, R, NS, RAD>(r * NS)) & 7);
                    }
                    if (slots.size() != 8) ++shared;
                }
        printf("N=%d T=%d %s NS=%d RAD=%d: %s, write groups sharing a slot: %d\n", N, T, sizeof(R) == 8 ? "fp64" : "fp32",
               NS, RAD, bad ? "NOT a linear bijection" : "linear bijection", shared);
        return bad + shared + check<N, T, R, NS * RAD>();
    }
}

int main()
{
    int bad = 0;
    bad += check<128, 64, double, 1>();
    bad += check<256, 64, double, 1>();
    bad += check<512, 128, double, 1>();
    bad += check<512, 64, double, 1>();
    bad += check<1024, 256, double, 1>();
    bad += check<1024, 128, double, 1>();
    bad += check<2048, 512, double, 1>();
    bad += check<2048, 256, double, 1>();
    bad += check<128, 64, float, 1>();
    bad += check<1024, 256, float, 1>();
    bad += check<1024, 128, float, 1>();
    bad += check<2048, 512, float, 1>();
    printf("failures: %d\n", bad);
    return bad != 0;
}
