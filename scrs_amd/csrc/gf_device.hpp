// gf_device.hpp - device-side GF(2^8) and GF((2^8)^2) arithmetic.
//
// GF(2^8): polynomial 0x11B, generator 3, log/exp tables staged in LDS by the
// caller (scrs/src/algebra/gf256.rs:29,39,261-300).  GF(2^16) is the tower
// a + b*u with u^2 = u + 0x20 (scrs/src/algebra/gf65536.rs:20,79-115), built
// on the same GF(2^8) tables: Karatsuba product, conjugate-over-norm inverse.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scrs {

struct Gf8Tables {
    const uint8_t *exp;  // 512 entries, doubled
    const uint8_t *log;  // 256 entries
    __device__ __forceinline__ uint32_t mul(uint32_t x, uint32_t y) const {
        return (x && y) ? exp[log[x] + log[y]] : 0u;
    }
    __device__ __forceinline__ uint32_t inv(uint32_t x) const { return x ? exp[255 - log[x]] : 0u; }
};

// Four packed GF(2^8) bytes times x (polynomial 0x11B), byte-wise.
__device__ __forceinline__ uint32_t gf8_xtime4(uint32_t v) {
    return ((v & 0x7f7f7f7fu) << 1) ^ (((v >> 7) & 0x01010101u) * 0x1bu);
}

// Field element operations over a uint32_t carrier.
struct FieldGf256 {
    Gf8Tables t;
    __device__ __forceinline__ uint32_t mul(uint32_t x, uint32_t y) const { return t.mul(x, y); }
    __device__ __forceinline__ uint32_t inv(uint32_t x) const { return t.inv(x); }
};

struct FieldGf65536 {
    Gf8Tables t;
    static constexpr uint32_t kDelta = 0x20;
    __device__ __forceinline__ uint32_t mul(uint32_t x, uint32_t y) const {
        const uint32_t a = x & 0xff, b = x >> 8, c = y & 0xff, d = y >> 8;
        const uint32_t ac = t.mul(a, c), bd = t.mul(b, d);
        const uint32_t lo = ac ^ t.mul(kDelta, bd);
        const uint32_t hi = t.mul(a ^ b, c ^ d) ^ ac;
        return lo | (hi << 8);
    }
    __device__ __forceinline__ uint32_t inv(uint32_t x) const {
        if (!x) return 0;
        const uint32_t a = x & 0xff, b = x >> 8;
        const uint32_t norm = t.mul(a, a) ^ t.mul(a, b) ^ t.mul(kDelta, t.mul(b, b));
        const uint32_t ni = t.inv(norm);
        return t.mul(a ^ b, ni) | (t.mul(b, ni) << 8);
    }
};

// Log-domain access for the closed-form solve, where every quantity is a
// product or quotient of coordinate differences: logs add, so long products
// become sums of independent lookups instead of chains of dependent
// multiplications.  ex() takes exponents below 2 * kOrder.
struct LogGf256 {
    Gf8Tables t;  // exp doubled (512), log (256) -- in LDS
    static constexpr uint32_t kOrder = 255;
    __device__ __forceinline__ uint32_t lg(uint32_t x) const { return t.log[x]; }
    __device__ __forceinline__ uint32_t ex(uint32_t e) const { return t.exp[e]; }
};

// GF(2^16) log / exp over the generator g = 0x0108 (order 65535,
// algebra/gf65536.rs:26): log[65536] then exp[2 * 65535], in device memory.
struct LogGf65536 {
    const uint16_t *tab;
    static constexpr uint32_t kOrder = 65535;
    __device__ __forceinline__ uint32_t lg(uint32_t x) const { return tab[x]; }
    __device__ __forceinline__ uint32_t ex(uint32_t e) const { return tab[65536u + e]; }
};

// Entries of the GF(2^16) log / exp table (log 65536, exp 2 * 65535).
constexpr uint32_t kGf16LogExpEntries = 65536u + 2u * 65535u;

// Host copy of the GF(2^8) exp (512, doubled) / log (256) tables.
void gf8_exp_log(uint8_t exp[512], uint8_t log[256]);

// Device copy of the GF(2^16) log / exp table for the current device, built
// on first use and kept until the process ends.
hipError_t gf16_log_exp_device(const uint16_t **out);

}  // namespace scrs
