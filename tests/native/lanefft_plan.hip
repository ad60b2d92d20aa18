// Host check of the lane-swap transform plans of cc_lanefft.hpp: for every
// (N, T) the kernels instantiate, prints the compile-time plan (digits, swap
// schedule, LDS stages) and the maps the kernels use, computed by the header's
// own functions, as one JSON object per line; exits non-zero if a map is not a
// bijection or an owner lookup disagrees with the spectral index map.
#include "../../parallel-critical-collapse_amd/csrc/cc_lanefft.hpp"
#include <cstdio>
#include <vector>

using namespace ccd;

template <int N, int T, int D>
static void pos_map(std::vector<int> &out)
{
    for (int t = 0; t < T; ++t)
        for (int r = 0; r < N / T; ++r) out.push_back(lf_pos_t<N, T, D>(t) | lf_pos_r<N, T, D>(r));
}

template <int N, int T, int D = 0>
static void all_pos_maps(std::vector<std::vector<int>> &maps)
{
    if constexpr (D < LF<N, T>::p.nd) {
        std::vector<int> m;
        pos_map<N, T, D>(m);
        maps.push_back(m);
        all_pos_maps<N, T, D + 1>(maps);
    }
}

static bool bijective(const std::vector<int> &v, int N)
{
    std::vector<char> seen(N, 0);
    for (int x : v) {
        if (x < 0 || x >= N || seen[x]) return false;
        seen[x] = 1;
    }
    return (int)v.size() == N;
}

template <int N, int T>
static int check()
{
    constexpr LfPlan p = LF<N, T>::p;
    constexpr int P = N / T;
    int bad = 0;
    std::vector<int> kmap;
    for (int t = 0; t < T; ++t)
        for (int r = 0; r < P; ++r) {
            const int k = lf_kt<N, T>(t) ^ lf_kr<N, T>(r);
            kmap.push_back(k);
            if (lf_owner<N, T>(k) != t * 64 + r) ++bad;
            if (lf_kr<N, T>(r) != T * r) ++bad;   // the register index is the top spectral digit
        }
    if (!bijective(kmap, N)) ++bad;
    std::vector<int> m64, m32;
    for (int k = 0; k < N; ++k) {
        m64.push_back(lf_mirror_pos<N, T, double>(k));
        m32.push_back(lf_mirror_pos<N, T, float>(k));
    }
    if (!bijective(m64, N) || !bijective(m32, N)) ++bad;
    std::vector<std::vector<int>> maps;
    all_pos_maps<N, T>(maps);
    for (auto &m : maps)
        if (!bijective(m, N)) ++bad;
    printf("{\"N\": %d, \"T\": %d, \"nd\": %d, \"nlds\": %d, \"bad\": %d, \"lo\": [", N, T, p.nd, p.nlds, bad);
    for (int d = 0; d < p.nd; ++d) printf("%s%d", d ? ", " : "", p.lo[d]);
    printf("], \"sz\": [");
    for (int d = 0; d < p.nd; ++d) printf("%s%d", d ? ", " : "", p.sz[d]);
    printf("], \"lds\": [");
    for (int d = 0; d < p.nd; ++d) printf("%s%d", d ? ", " : "", p.lds[d]);
    printf("], \"sw\": [");
    for (int d = 0; d < p.nd; ++d) {
        printf("%s[", d ? ", " : "");
        for (int m = 0; d > 0 && m < p.sz[d]; ++m) printf("%s%d", m ? ", " : "", p.sw[d][m]);
        printf("]");
    }
    printf("], \"k\": [");
    for (size_t i = 0; i < kmap.size(); ++i) printf("%s%d", i ? ", " : "", kmap[i]);
    printf("], \"mirror64\": [");
    for (size_t i = 0; i < m64.size(); ++i) printf("%s%d", i ? ", " : "", m64[i]);
    printf("]}\n");
    return bad;
}

int main()
{
    int bad = 0;
    bad += check<128, 64>();
    bad += check<256, 64>();
    bad += check<512, 128>();
    bad += check<512, 64>();
    bad += check<1024, 256>();
    bad += check<1024, 128>();
    bad += check<2048, 512>();
    bad += check<2048, 256>();
    return bad ? 1 : 0;
}
