// field.hpp - host-side GF(2^8) and GF((2^8)^2) arithmetic for coefficient
// generation and the small per-codeword solves done on the host.
//
// GF(2^8) is the AES field, polynomial 0x11B, generator 0x03
// (scrs/src/algebra/gf256.rs:29,39).  GF(2^16) is the quadratic extension
// a + b*u with u^2 + u + 0x20 = 0, stored little-endian as [a, b]
// (scrs/src/algebra/gf65536.rs:1-20,53-60).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>

namespace scrs {

struct Gf256 {
    std::array<uint8_t, 512> exp{};  // exp[i] = 3^i, doubled so log sums need no modulo
    std::array<uint8_t, 256> log{};  // log[3^i] = i, log[0] unused
    std::array<std::array<uint8_t, 256>, 256> mul{};

    Gf256() {
        uint8_t v = 1;
        for (int i = 0; i < 255; i++) {
            exp[i] = v;
            exp[i + 255] = v;
            log[v] = static_cast<uint8_t>(i);
            // v *= 3  ->  v ^ xtime(v)
            uint8_t x2 = static_cast<uint8_t>((v << 1) ^ ((v & 0x80) ? 0x1B : 0));
            v = static_cast<uint8_t>(v ^ x2);
        }
        exp[510] = exp[0];
        exp[511] = exp[1];
        for (int a = 0; a < 256; a++)
            for (int b = 0; b < 256; b++)
                mul[a][b] = (a && b) ? exp[log[a] + log[b]] : 0;
    }
    uint8_t m(uint8_t a, uint8_t b) const { return mul[a][b]; }
    uint8_t inv(uint8_t a) const { return a ? exp[255 - log[a]] : 0; }
    uint8_t pow3(size_t e) const { return exp[e % 255]; }
};

const Gf256 &gf256();

// GF((2^8)^2) element operations (algebra/gf65536.rs:79-144).
struct Gf16 {
    static constexpr uint8_t kDelta = 0x20;
    static constexpr uint16_t kGenerator = 0x0108;

    static uint16_t mul(uint16_t x, uint16_t y) {
        const Gf256 &f = gf256();
        uint8_t a = static_cast<uint8_t>(x), b = static_cast<uint8_t>(x >> 8);
        uint8_t c = static_cast<uint8_t>(y), d = static_cast<uint8_t>(y >> 8);
        uint8_t ac = f.m(a, c), bd = f.m(b, d);
        uint8_t lo = ac ^ f.m(kDelta, bd);
        uint8_t hi = f.m(a ^ b, c ^ d) ^ ac;
        return static_cast<uint16_t>(lo | (hi << 8));
    }
    static uint16_t inv(uint16_t x) {
        if (!x) return 0;
        const Gf256 &f = gf256();
        uint8_t a = static_cast<uint8_t>(x), b = static_cast<uint8_t>(x >> 8);
        uint8_t norm = f.m(a, a) ^ f.m(a, b) ^ f.m(kDelta, f.m(b, b));
        uint8_t ni = f.inv(norm);
        return static_cast<uint16_t>(f.m(a ^ b, ni) | (f.m(b, ni) << 8));
    }
};

}  // namespace scrs
