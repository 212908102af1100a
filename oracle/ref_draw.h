// ref_draw.h -- the reference annealer's draws, for the oracle's models.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). One worker of the reference draws
// from one std::mt19937 stream (congruence_anneal.hpp:189-192): libstdc++'s
// uniform_int_distribution over active positions and
// uniform_real_distribution(0, 1), consumed in the loop's order, with the
// redraw loops for j != i (:287) and for the fallback pivot (:226), the sign
// from the low bit of a raw draw (:293, :295) and std::exp(-d / T) in the
// Metropolis test (:236). The models' loop bodies take this or the engine's
// Philox words through the same calls.
#pragma once

#include <cmath>
#include <cstdint>
#include <random>

struct RefDraw {
    std::mt19937 rng;
    std::uniform_int_distribution<int> aidx;
    std::uniform_real_distribution<double> unit{0.0, 1.0};
    RefDraw(uint32_t seed, int N) : rng(seed), aidx(0, N - 1) {}
    bool coin(double f) { return unit(rng) < f; }
    uint32_t index() { return (uint32_t)aidx(rng); }
    uint32_t pivot_index() { return (uint32_t)aidx(rng); }
    uint32_t fallback(uint32_t tpos) {
        uint32_t q;
        do q = (uint32_t)aidx(rng);
        while (q == tpos);
        return q;
    }
    void pair(uint32_t& ip, uint32_t& jp) {
        ip = (uint32_t)aidx(rng);
        do jp = (uint32_t)aidx(rng);
        while (jp == ip);
    }
    bool greedy_coin(double f) { return unit(rng) < f; }
    int64_t sign() { return (rng() & 1u) ? 1 : -1; }
    // d: the move's score change as a double (an integer for the L1 score)
    template <class Delta>
    bool metropolis(Delta d, double T) {
        return unit(rng) < std::exp(-(double)d / T);
    }
    bool sweep_coin(double f) { return unit(rng) < f; }
    uint32_t sweep_index() { return (uint32_t)aidx(rng); }
};
