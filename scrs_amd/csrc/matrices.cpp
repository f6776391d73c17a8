// matrices.cpp - coefficient matrices and packed product tables.
#include "matrices.hpp"

namespace scrs {

const Gf256 &gf256() {
    static const Gf256 field;
    return field;
}

std::vector<uint8_t> standard_cauchy_matrix(size_t k, size_t m) {
    const Gf256 &f = gf256();
    std::vector<uint8_t> c(k * m);
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++)
            c[i * m + j] = f.inv(static_cast<uint8_t>(i ^ (k + j)));
    return c;
}

std::vector<uint8_t> good_cauchy_matrix(size_t k, size_t m) {
    const Gf256 &f = gf256();
    std::vector<uint8_t> base(k + m, 0);
    for (size_t d = 1; d < k + m; d++) base[d] = f.inv(static_cast<uint8_t>(1 ^ f.pow3(d)));
    std::vector<uint8_t> c(k * m);
    for (size_t i = 0; i < k; i++) {
        uint8_t scale = f.inv(f.pow3(i));
        for (size_t j = 0; j < m; j++) c[i * m + j] = f.m(scale, base[k + j - i]);
    }
    return c;
}

std::vector<uint16_t> tower_cauchy_matrix(size_t k, size_t m) {
    const size_t n = k + m;
    std::vector<uint16_t> powers(n);
    uint16_t v = 1;
    for (size_t i = 0; i < n; i++) {
        powers[i] = v;
        v = Gf16::mul(v, Gf16::kGenerator);
    }
    // base[d] = 1 / (1 + g^d) for d = 1..n-1, inverted together with one
    // field inversion (prefix products, then one reciprocal walked back).
    std::vector<uint16_t> base(n, 0), prefix(n, 1);
    uint16_t product = 1;
    for (size_t d = 1; d < n; d++) {
        prefix[d] = product;
        product = Gf16::mul(product, static_cast<uint16_t>(1 ^ powers[d]));
    }
    uint16_t r = Gf16::inv(product);
    for (size_t d = n - 1; d >= 1; d--) {
        uint16_t value = static_cast<uint16_t>(1 ^ powers[d]);
        base[d] = Gf16::mul(r, prefix[d]);
        r = Gf16::mul(r, value);
    }
    const uint16_t ginv = Gf16::inv(Gf16::kGenerator);
    std::vector<uint16_t> c(k * m);
    uint16_t scale = 1;
    for (size_t i = 0; i < k; i++) {
        for (size_t j = 0; j < m; j++) c[i * m + j] = Gf16::mul(scale, base[k + j - i]);
        scale = Gf16::mul(scale, ginv);
    }
    return c;
}

void build_nibble_tables(const uint8_t *coef, size_t T, size_t R, size_t groups,
                         uint32_t *out) {
    const Gf256 &f = gf256();
    for (size_t t = 0; t < T; t++)
        for (size_t g = 0; g < groups; g++) {
            uint8_t *tab = reinterpret_cast<uint8_t *>(out + (t * groups + g) * 128);
            for (int n = 0; n < 16; n++)
                for (int s = 0; s < 16; s++) {
                    const size_t r = g * 16 + s;
                    const uint8_t c = r < R ? coef[t * R + r] : 0;
                    tab[n * 16 + s] = f.m(c, static_cast<uint8_t>(n));
                    tab[(16 + n) * 16 + s] = f.m(c, static_cast<uint8_t>(n << 4));
                }
        }
}

void build_nibble_tables16(const uint16_t *coef, size_t T, size_t R, size_t groups,
                           uint32_t *out) {
    for (size_t t = 0; t < T; t++)
        for (size_t g = 0; g < groups; g++) {
            uint8_t *tab = reinterpret_cast<uint8_t *>(out + (t * groups + g) * 512);
            for (int s = 0; s < 16; s++) {
                const size_t r = g * 16 + s;
                const uint16_t c = r < R ? coef[t * R + r] : 0;
                uint16_t basis[16];
                for (int b = 0; b < 16; b++) basis[b] = Gf16::mul(c, static_cast<uint16_t>(1u << b));
                for (int q = 0; q < 4; q++)
                    for (int n = 0; n < 16; n++) {
                        uint16_t v = 0;
                        for (int b = 0; b < 4; b++)
                            if ((n >> b) & 1) v ^= basis[4 * q + b];
                        uint8_t *e = tab + (q * 16 + n) * 32 + 2 * s;
                        e[0] = static_cast<uint8_t>(v);
                        e[1] = static_cast<uint8_t>(v >> 8);
                    }
            }
        }
}

}  // namespace scrs
