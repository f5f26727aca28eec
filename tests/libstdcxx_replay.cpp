// Feeds given 32-bit words to libstdc++'s own std::discrete_distribution and
// std::generate_canonical through a replaying generator, so that a test can
// compare the oracle's draw -> index mapping with the library the reference
// links (cp_arls_lev.hpp:106,145 and sts_cp.hpp:198 draw through them).
//
// stdin:  W w_0 .. w_{W-1}  D  x0_0 x1_0 .. x0_{D-1} x1_{D-1}
// stdout: per draw "<index> <canonical as %a>"; the canonical value is
// generate_canonical<double, 53> of the same two words, the index is the
// discrete_distribution draw from those two words.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

struct Replay {
    typedef uint32_t result_type;
    const std::vector<uint32_t>* words;
    size_t at = 0;
    static constexpr result_type min() { return 0u; }
    static constexpr result_type max() { return 0xFFFFFFFFu; }
    result_type operator()() { return (*words)[at++]; }
};

int main() {
    size_t W = 0, D = 0;
    if (std::scanf("%zu", &W) != 1) return 1;
    std::vector<double> w(W);
    for (auto& x : w)
        if (std::scanf("%lf", &x) != 1) return 1;
    if (std::scanf("%zu", &D) != 1) return 1;
    std::vector<uint32_t> words(2 * D);
    for (auto& x : words)
        if (std::scanf("%u", &x) != 1) return 1;
    std::discrete_distribution<int> dist(w.begin(), w.end());
    for (size_t i = 0; i < D; i++) {
        std::vector<uint32_t> two{words[2 * i], words[2 * i + 1]};
        Replay a{&two};
        const int idx = dist(a);
        Replay b{&two};
        const double u = std::generate_canonical<double, 53>(b);
        std::printf("%d %a %zu\n", idx, u, a.at);
    }
    return 0;
}
