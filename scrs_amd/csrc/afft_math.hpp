// afft_math.hpp - host-side constants of the additive FFT over GF(2^16).
//
// The transform evaluates polynomials in the Lin-Chung-Han novel basis at
// the points 0..N-1 (N = 2^D <= 65536).  Its butterflies use one factor per
// node of a binary tree; a node is identified by its dimension d (it spans
// 2^d consecutive points) and its shift s (the first point, a multiple of
// 2^d), and its factor is the normalised subspace polynomial
// W_{d-1}(s) / W_{d-1}(2^(d-1)) (scrs/src/afft/transform.rs:229-299).
// Factors depend only on (d, s), so every plan size shares one table laid
// out as the 65536-point heap: node(d, s) = 2^(16-d) + (s >> d).
#pragma once

#include <cstdint>
#include <vector>

namespace scrs {

constexpr uint32_t kAfftMaxLog = 16;
constexpr uint32_t kAfftOrder = 65535;  // multiplicative group order of GF(2^16)

struct AfftMath {
    std::vector<uint16_t> exp;      // exp[i] = g^i, i in [0, 65535), g = 0x0108
    std::vector<uint16_t> log;      // log[g^i] = i; log[0] = 0 (a product skips a zero factor)
    std::vector<uint16_t> factors;  // 65536 entries, heap order, entry 0 unused
    uint16_t deriv[kAfftMaxLog];    // derivative factor of bit b (transform.rs:45-49,188-215)
};

const AfftMath &afft_math();

inline uint32_t afft_node(uint32_t dim, uint32_t shift) { return (1u << (kAfftMaxLog - dim)) + (shift >> dim); }

// Walsh-Hadamard transform mod 65535 of log[x] for x in [0, 2^D).
std::vector<uint32_t> afft_log_wht(uint32_t D);

// Erasure-locator values over the domain [0, 2^D) for the point set `known`
// (afft/decoder.rs:672-714): for a known x, c[x] = prod over unknown e of
// (x + e); for an unknown x, c[x] = 1 / prod over unknown e != x of (x + e).
std::vector<uint16_t> afft_locator_scales(const std::vector<uint8_t> &known, uint32_t D);

uint16_t gf16_mul_host(uint16_t a, uint16_t b);

}  // namespace scrs
