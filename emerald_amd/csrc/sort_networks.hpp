// Sorting networks as compile-time comparator lists (plain C++: the device
// code unrolls them, and a host program checks each one by the 0-1 principle).
//
//   8 inputs:  19 comparators, depth 6 (optimal size)
//   12 inputs: 39 comparators, depth 9 (optimal size)
//   16, 32, 64 inputs: Batcher's odd-even merge network (63, 191, 543)
//
// The 8- and 12-input networks are the published minimal-size networks
// (Knuth, TAOCP vol. 3, section 5.3.4).
#ifndef EMSPH_SORT_NETWORKS_HPP
#define EMSPH_SORT_NETWORKS_HPP

namespace emsph {

struct Comparator {
    int a;
    int b;
};

template <int N, int M>
struct FixedNet {
    static constexpr int inputs = N;
    Comparator ce[M];
    int count;
};

constexpr FixedNet<8, 19> NET8{{{0, 2}, {1, 3}, {4, 6}, {5, 7},
                                {0, 4}, {1, 5}, {2, 6}, {3, 7},
                                {0, 1}, {2, 3}, {4, 5}, {6, 7},
                                {2, 4}, {3, 5},
                                {1, 4}, {3, 6},
                                {1, 2}, {3, 4}, {5, 6}},
                               19};

constexpr FixedNet<12, 39> NET12{{{0, 1}, {2, 3}, {4, 5}, {6, 7}, {8, 9}, {10, 11},
                                  {1, 3}, {5, 7}, {9, 11}, {0, 2}, {4, 6}, {8, 10},
                                  {1, 2}, {5, 6}, {9, 10}, {0, 4}, {7, 11},
                                  {1, 5}, {6, 10}, {3, 7}, {4, 8},
                                  {5, 9}, {2, 6}, {0, 4}, {7, 11}, {3, 8},
                                  {1, 5}, {6, 10}, {2, 3}, {8, 9},
                                  {1, 4}, {7, 10}, {3, 5}, {6, 8},
                                  {2, 4}, {7, 9}, {5, 6},
                                  {3, 4}, {7, 8}},
                                 39};

// Batcher's odd-even merge sorting network on N inputs (N a power of two).
template <int N>
struct OddEvenMergeNet {
    static constexpr int inputs = N;
    Comparator ce[N * N];
    int count = 0;
    constexpr OddEvenMergeNet() : ce{} {
        for (int p = 1; p < N; p <<= 1) {
            for (int k = p; k >= 1; k >>= 1) {
                for (int j = k % p; j <= N - 1 - k; j += 2 * k) {
                    int const last = (k - 1 < N - j - k - 1) ? k - 1 : N - j - k - 1;
                    for (int i = 0; i <= last; ++i) {
                        if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                            ce[count].a = i + j;
                            ce[count].b = i + j + k;
                            ++count;
                        }
                    }
                }
            }
        }
    }
};
constexpr OddEvenMergeNet<16> NET16{};
constexpr OddEvenMergeNet<32> NET32{};
constexpr OddEvenMergeNet<64> NET64{};

}  // namespace emsph

#endif  // EMSPH_SORT_NETWORKS_HPP
