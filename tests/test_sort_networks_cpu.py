"""The neighbour build's sorting networks (emerald_amd/csrc/sort_networks.hpp)
sort every input: by the 0-1 principle, a comparator network sorts all inputs
iff it sorts every sequence of 0s and 1s. The 8-, 12- and 16-input networks
are checked on all 2^N 0-1 inputs, the 32- and 64-input ones on random 0-1
inputs and random permutations. A network with one comparator removed must
fail the same check."""

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "emerald_amd", "csrc", "sort_networks.hpp")

CHECK = r'''
#include <cstdint>
#include <cstdio>
#include <random>
#include <algorithm>
#include "sort_networks.hpp"
using namespace emsph;
template <class Net>
bool sorts(Net const& net, uint64_t bits, int n, int drop) {
    int v[64];
    for (int i = 0; i < n; ++i) v[i] = (bits >> i) & 1;
    for (int c = 0; c < net.count; ++c) {
        if (c == drop) continue;
        int a = net.ce[c].a, b = net.ce[c].b;
        if (v[a] > v[b]) std::swap(v[a], v[b]);
    }
    for (int i = 0; i + 1 < n; ++i) if (v[i] > v[i + 1]) return false;
    return true;
}
template <class Net>
bool perm_sorts(Net const& net, int n, std::mt19937_64& g) {
    int v[64];
    for (int i = 0; i < n; ++i) v[i] = i;
    std::shuffle(v, v + n, g);
    for (int c = 0; c < net.count; ++c) {
        int a = net.ce[c].a, b = net.ce[c].b;
        if (v[a] > v[b]) std::swap(v[a], v[b]);
    }
    for (int i = 0; i < n; ++i) if (v[i] != i) return false;
    return true;
}
template <class Net>
int check(Net const& net, int n, int drop) {
    if (n <= 16) {
        for (uint64_t m = 0; m < (1ull << n); ++m) if (!sorts(net, m, n, drop)) return 1;
        return 0;
    }
    std::mt19937_64 g(12345);
    for (int k = 0; k < 200000; ++k) if (!sorts(net, g(), n, drop)) return 1;
    for (int k = 0; k < 20000; ++k) if (drop < 0 && !perm_sorts(net, n, g)) return 1;
    return 0;
}
int main() {
    int bad = 0;
    bad |= check(NET8, 8, -1) << 0;
    bad |= check(NET12, 12, -1) << 1;
    bad |= check(NET16, 16, -1) << 2;
    bad |= check(NET32, 32, -1) << 3;
    bad |= check(NET64, 64, -1) << 4;
    // Removing any single comparator of the minimal networks breaks them.
    for (int c = 0; c < NET8.count; ++c) if (!check(NET8, 8, c)) bad |= 1 << 5;
    for (int c = 0; c < NET12.count; ++c) if (!check(NET12, 12, c)) bad |= 1 << 6;
    std::printf("%d %d %d %d %d %d\n", NET8.count, NET12.count, NET16.count, NET32.count,
                NET64.count, bad);
    return bad;
}
'''


def test_sorting_networks_sort_every_input(tmp_path):
    src = tmp_path / "check.cpp"
    src.write_text(CHECK)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.dirname(HEADER), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    counts = [int(x) for x in out.stdout.split()]
    assert counts[:5] == [19, 39, 63, 191, 543]
