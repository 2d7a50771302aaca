// rng.hpp -- xoshiro256++ seeded through SplitMix64 (Blackman & Vigna), the
// generator family the reference's SmallRng uses on 64-bit targets
// (partition/kway.rs:62 seeds it with seed_from_u64).  Used by the host for
// per-level matching seeds and the BFS growing of the coarsest graph.
#pragma once
#include <cstdint>

namespace rk {

struct Rng {
    uint64_t s[4];

    static uint64_t splitmix(uint64_t& st) {
        st += 0x9e3779b97f4a7c15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t st = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix(st);
        if (!(s[0] | s[1] | s[2] | s[3])) s[0] = 1;
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next_u64() {
        const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // Uniform integer in [0, n) by widening multiply with rejection.
    uint64_t below(uint64_t n) {
        if (n == 0) return next_u64();
        const uint64_t zone = (n << __builtin_clzll(n)) - 1;
        for (;;) {
            const unsigned __int128 p = (unsigned __int128)next_u64() * n;
            if ((uint64_t)p <= zone) return (uint64_t)(p >> 64);
        }
    }
};

}  // namespace rk
