// afft_math.cpp - host-side constants of the additive FFT over GF(2^16).
#include "afft_math.hpp"

#include <mutex>

#include "field.hpp"

namespace scrs {

uint16_t gf16_mul_host(uint16_t a, uint16_t b) { return Gf16::mul(a, b); }

namespace {

uint16_t gf16_square(uint16_t a) { return Gf16::mul(a, a); }

// Linearised polynomial sum_j c[j] * x^(2^j).
uint16_t eval_linearized(const std::vector<uint16_t> &c, uint16_t x) {
    uint16_t acc = 0, p = x;
    for (uint16_t cj : c) {
        acc ^= Gf16::mul(cj, p);
        p = gf16_square(p);
    }
    return acc;
}

AfftMath *build() {
    auto *m = new AfftMath();
    m->exp.resize(kAfftOrder);
    m->log.assign(65536, 0);
    uint16_t v = 1;
    for (uint32_t i = 0; i < kAfftOrder; i++) {
        m->exp[i] = v;
        m->log[v] = static_cast<uint16_t>(i);
        v = Gf16::mul(v, Gf16::kGenerator);
    }
    // Subspace polynomials W_0(x) = x, W_{i+1} = W_i^2 + W_i(2^i) W_i, kept
    // as linearised coefficients; each is normalised by W_i(2^i).
    std::vector<std::vector<uint16_t>> poly;
    std::vector<uint16_t> norm_inv;
    std::vector<uint16_t> c{1};
    for (uint32_t i = 0; i < kAfftMaxLog; i++) {
        const uint16_t nv = eval_linearized(c, static_cast<uint16_t>(1u << i));
        poly.push_back(c);
        norm_inv.push_back(Gf16::inv(nv));
        m->deriv[i] = Gf16::mul(c[0], norm_inv.back());
        std::vector<uint16_t> next(c.size() + 1, 0);
        for (size_t j = 0; j < c.size(); j++) {
            next[j] ^= Gf16::mul(nv, c[j]);
            next[j + 1] ^= gf16_square(c[j]);
        }
        c = next;
    }
    // Node factors.  The normalised polynomial is GF(2)-linear, so its value
    // at a shift is the XOR of its values at the shift's set bits.
    m->factors.assign(65536, 0);
    for (uint32_t d = 1; d <= kAfftMaxLog; d++) {
        uint16_t bv[kAfftMaxLog] = {};
        for (uint32_t j = d; j < kAfftMaxLog; j++)
            bv[j] = Gf16::mul(eval_linearized(poly[d - 1], static_cast<uint16_t>(1u << j)), norm_inv[d - 1]);
        const uint32_t count = 1u << (kAfftMaxLog - d);
        std::vector<uint16_t> f(count, 0);
        for (uint32_t q = 1; q < count; q++) {
            const uint32_t low = __builtin_ctz(q);
            f[q] = f[q & (q - 1)] ^ bv[d + low];
        }
        for (uint32_t q = 0; q < count; q++) m->factors[count + q] = f[q];
    }
    return m;
}

void wht_mod(std::vector<uint32_t> &v) {
    const size_t n = v.size();
    for (size_t h = 1; h < n; h <<= 1)
        for (size_t b = 0; b < n; b += 2 * h)
            for (size_t i = b; i < b + h; i++) {
                const uint32_t x = v[i], y = v[i + h];
                const uint32_t s = x + y;
                v[i] = s >= kAfftOrder ? s - kAfftOrder : s;
                v[i + h] = x >= y ? x - y : x + kAfftOrder - y;
            }
}

uint32_t mulmod(uint32_t a, uint32_t b) { return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) % kAfftOrder); }

}  // namespace

const AfftMath &afft_math() {
    static std::once_flag once;
    static AfftMath *m = nullptr;
    std::call_once(once, [] { m = build(); });
    return *m;
}

std::vector<uint32_t> afft_log_wht(uint32_t D) {
    const AfftMath &m = afft_math();
    const size_t n = size_t{1} << D;
    std::vector<uint32_t> v(n);
    for (size_t x = 0; x < n; x++) v[x] = m.log[x];
    wht_mod(v);
    return v;
}

std::vector<uint16_t> afft_locator_scales(const std::vector<uint8_t> &known, uint32_t D) {
    const AfftMath &m = afft_math();
    const size_t n = size_t{1} << D;
    // log of the locator at x is the XOR-convolution of the unknown-point
    // indicator with log(.), computed through three transforms mod 65535.
    std::vector<uint32_t> ind(n);
    for (size_t x = 0; x < n; x++) ind[x] = known[x] ? 0u : 1u;
    wht_mod(ind);
    const std::vector<uint32_t> lw = afft_log_wht(D);
    for (size_t x = 0; x < n; x++) ind[x] = mulmod(ind[x], lw[x]);
    wht_mod(ind);
    uint32_t inv_n = 1;  // 2^-D mod 65535 = 32768^D
    for (uint32_t i = 0; i < D; i++) inv_n = mulmod(inv_n, 32768);
    std::vector<uint16_t> c(n);
    for (size_t x = 0; x < n; x++) {
        const uint32_t e = mulmod(ind[x], inv_n);
        c[x] = known[x] ? m.exp[e] : m.exp[(kAfftOrder - e) % kAfftOrder];
    }
    return c;
}

}  // namespace scrs
