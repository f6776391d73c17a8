/*
 * scrs_oracle.c - CPU restatement of the reference's payload arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  Product code never links or calls it.
 *
 * Citations are reference file:line, relative to the reference repository.
 */
#include "scrs_oracle.h"
#include "../include/scrs_status.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== */
/* GF(2^8)                                                                 */
/* ====================================================================== */

/* Russian-peasant multiply with xtime reduction by 0x1B
 * (algebra/gf256.rs:79-98, REDUCTION_POLY 0x11B at :29). */
uint8_t or_gf256_mul(uint8_t a, uint8_t b) {
    uint8_t acc = 0;
    for (int i = 0; i < 8; i++) {
        if ((b >> i) & 1) acc ^= a;
        uint8_t hi = a & 0x80;
        a = (uint8_t)(a << 1);
        if (hi) a ^= 0x1B;
    }
    return acc;
}

/* Fermat inverse a^254, zero maps to zero (algebra/gf256.rs:105-127). */
uint8_t or_gf256_inv(uint8_t a) {
    if (a == 0) return 0;
    uint8_t r = 1;
    for (int i = 0; i < 8; i++) {
        r = or_gf256_mul(r, r);
        if ((254u >> (7 - i)) & 1) r = or_gf256_mul(r, a);
    }
    return r;
}

/* EXP[i] = 3^i, LOG inverts it, LOG[0] = 0 (algebra/gf256.rs:261-300). */
void or_gf256_exp_log(uint8_t exp_out[255], uint8_t log_out[256]) {
    uint8_t v = 1;
    memset(log_out, 0, 256);
    for (int i = 0; i < 255; i++) {
        exp_out[i] = v;
        log_out[v] = (uint8_t)i;
        v = or_gf256_mul(v, 0x03);
    }
}

#if defined(__x86_64__)
#include <immintrin.h>
__attribute__((target("gfni,sse2"))) static uint8_t gfni_mul1(uint8_t a, uint8_t b) {
    __m128i x = _mm_set1_epi8((char)a), y = _mm_set1_epi8((char)b);
    return (uint8_t)_mm_cvtsi128_si32(_mm_gf2p8mul_epi8(x, y));
}
#endif

/* The host CPU's GF2P8MULB multiplies in the same 0x11B field; the
 * reference leans on the same fact (simd/x86.rs:1402-1425). */
int or_gf256_mul_gfni(uint8_t a, uint8_t b) {
#if defined(__x86_64__)
    __builtin_cpu_init();
    if (!__builtin_cpu_supports("gfni")) return -1;
    return gfni_mul1(a, b);
#else
    (void)a; (void)b;
    return -1;
#endif
}

/* ====================================================================== */
/* GF((2^8)^2)                                                             */
/* ====================================================================== */

#define GF16_DELTA 0x20 /* algebra/gf65536.rs:20 */

/* Karatsuba product [ac + d*bd, (a+b)(c+d) + ac] (algebra/gf65536.rs:79-90). */
uint16_t or_gf16_mul(uint16_t x, uint16_t y) {
    uint8_t a = (uint8_t)x, b = (uint8_t)(x >> 8);
    uint8_t c = (uint8_t)y, d = (uint8_t)(y >> 8);
    uint8_t ac = or_gf256_mul(a, c), bd = or_gf256_mul(b, d);
    uint8_t lo = ac ^ or_gf256_mul(GF16_DELTA, bd);
    uint8_t hi = or_gf256_mul(a ^ b, c ^ d) ^ ac;
    return (uint16_t)(lo | (hi << 8));
}

/* Schoolbook form used by the reference's own cross-check
 * (algebra/gf65536.rs:202-209). */
uint16_t or_gf16_mul_schoolbook(uint16_t x, uint16_t y) {
    uint8_t a = (uint8_t)x, b = (uint8_t)(x >> 8);
    uint8_t c = (uint8_t)y, d = (uint8_t)(y >> 8);
    uint8_t lo = or_gf256_mul(a, c) ^ or_gf256_mul(GF16_DELTA, or_gf256_mul(b, d));
    uint8_t hi = or_gf256_mul(a, d) ^ or_gf256_mul(b, c) ^ or_gf256_mul(b, d);
    return (uint16_t)(lo | (hi << 8));
}

/* (a + bu)^2 = a^2 + d b^2 + b^2 u (algebra/gf65536.rs:93-99). */
uint16_t or_gf16_square(uint16_t x) {
    uint8_t a = (uint8_t)x, b = (uint8_t)(x >> 8);
    uint8_t a2 = or_gf256_mul(a, a), b2 = or_gf256_mul(b, b);
    return (uint16_t)((a2 ^ or_gf256_mul(GF16_DELTA, b2)) | (b2 << 8));
}

/* Conjugate over norm (algebra/gf65536.rs:104-115). */
uint16_t or_gf16_inv(uint16_t x) {
    uint8_t a = (uint8_t)x, b = (uint8_t)(x >> 8);
    if (x == 0) return 0;
    uint8_t norm = or_gf256_mul(a, a) ^ or_gf256_mul(a, b) ^
                   or_gf256_mul(GF16_DELTA, or_gf256_mul(b, b));
    uint8_t ni = or_gf256_inv(norm);
    return (uint16_t)(or_gf256_mul(a ^ b, ni) | (or_gf256_mul(b, ni) << 8));
}

/* Square-and-multiply (algebra/gf65536.rs:132-144). */
uint16_t or_gf16_pow(uint16_t base, uint32_t e) {
    uint16_t r = 1;
    while (e) {
        if (e & 1) r = or_gf16_mul(r, base);
        base = or_gf16_square(base);
        e >>= 1;
    }
    return r;
}

/* ====================================================================== */
/* Coefficient matrices                                                    */
/* ====================================================================== */

/* Standard Cauchy: X = {0..k-1}, Y = {k..k+m-1}, C = 1/(x+y), capacity 256
 * (matrices/cauchy.rs:38-49, :70-75, :88-99). */
size_t or_std_cauchy_matrix(size_t k, size_t m, uint8_t *out) {
    if (k + m > 256) return 256;
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++)
            out[i * m + j] = or_gf256_inv((uint8_t)(i ^ (k + j)));
    return 0;
}

static uint8_t gexp(size_t e) {
    uint8_t r = 1;
    for (size_t i = 0; i < e % 255; i++) r = or_gf256_mul(r, 0x03);
    return r;
}

/* Good Cauchy, direct form: X = {3^i}, Y = {3^(k+j)}, capacity 255
 * (matrices/good_cauchy.rs:59-64, :94-101, :211-231). */
size_t or_good_cauchy_matrix(size_t k, size_t m, uint8_t *out) {
    if (k + m > 255) return 255;
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++)
            out[i * m + j] = or_gf256_inv(gexp(i) ^ gexp(k + j));
    return 0;
}

/* Good Cauchy via the diagonal factorisation 3^-i / (1 + 3^(k+j-i)),
 * the form the reference materialises (matrices/good_cauchy.rs:125-151). */
size_t or_good_cauchy_matrix_factored(size_t k, size_t m, uint8_t *out) {
    if (k + m > 255) return 255;
    for (size_t i = 0; i < k; i++) {
        uint8_t scale = i == 0 ? 1 : or_gf256_inv(gexp(i));
        for (size_t j = 0; j < m; j++) {
            size_t d = k + j - i;
            out[i * m + j] = or_gf256_mul(scale, or_gf256_inv((uint8_t)(1 ^ gexp(d))));
        }
    }
    return 0;
}

/* Tower Cauchy over GF(2^16): X = {g^i}, Y = {g^(k+j)}, g = 0x0108,
 * capacity 65535 (tower/cauchy.rs:6, :46-63; algebra/gf65536.rs:26). */
size_t or_tower_cauchy_matrix(size_t k, size_t m, uint16_t *out) {
    if (k + m > 65535) return 65535;
    size_t n = k + m;
    uint16_t *pw = (uint16_t *)malloc(n * sizeof(uint16_t));
    uint16_t v = 1;
    for (size_t i = 0; i < n; i++) {
        pw[i] = v;
        v = or_gf16_mul(v, 0x0108);
    }
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++)
            out[i * m + j] = or_gf16_inv(pw[i] ^ pw[k + j]);
    free(pw);
    return 0;
}

/* ====================================================================== */
/* GF(256) codec                                                           */
/* ====================================================================== */

/* Repair-major reference encode (batch/codec.rs:199-215,
 * encode_into_reference). */
void or_gf256_encode(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                     const uint8_t *data, uint8_t *repairs) {
    for (size_t j = 0; j < m; j++) {
        uint8_t *r = repairs + j * L;
        memset(r, 0, L);
        for (size_t i = 0; i < k; i++) {
            uint8_t c = coeffs[i * m + j];
            const uint8_t *s = data + i * L;
            for (size_t p = 0; p < L; p++) r[p] ^= or_gf256_mul(c, s[p]);
        }
    }
}

void or_gf256_rows_terms(uint8_t *dst, size_t L, size_t rows, size_t terms,
                         const uint8_t *coef, const uint8_t *const *srcs) {
    for (size_t t = 0; t < terms; t++)
        for (size_t r = 0; r < rows; r++) {
            uint8_t c = coef[t * rows + r];
            for (size_t p = 0; p < L; p++) dst[r * L + p] ^= or_gf256_mul(c, srcs[t][p]);
        }
}

/* Gauss-Jordan with first-nonzero pivoting, inverse written row-major
 * (batch/codec.rs:525-562, invert_square_into).  Returns 0 if singular. */
static int gf256_invert(uint8_t *mat, size_t e, uint8_t *inv) {
    memset(inv, 0, e * e);
    for (size_t i = 0; i < e; i++) inv[i * e + i] = 1;
    for (size_t col = 0; col < e; col++) {
        size_t piv = col;
        while (piv < e && mat[piv * e + col] == 0) piv++;
        if (piv == e) return 0;
        if (piv != col)
            for (size_t c = 0; c < e; c++) {
                uint8_t t = mat[col * e + c]; mat[col * e + c] = mat[piv * e + c]; mat[piv * e + c] = t;
                t = inv[col * e + c]; inv[col * e + c] = inv[piv * e + c]; inv[piv * e + c] = t;
            }
        uint8_t s = or_gf256_inv(mat[col * e + col]);
        for (size_t c = 0; c < e; c++) {
            mat[col * e + c] = or_gf256_mul(mat[col * e + c], s);
            inv[col * e + c] = or_gf256_mul(inv[col * e + c], s);
        }
        for (size_t r = 0; r < e; r++) {
            if (r == col) continue;
            uint8_t f = mat[r * e + col];
            if (!f) continue;
            for (size_t c = 0; c < e; c++) {
                mat[r * e + c] ^= or_gf256_mul(f, mat[col * e + c]);
                inv[r * e + c] ^= or_gf256_mul(f, inv[col * e + c]);
            }
        }
    }
    return 1;
}

#define RET(code_, a_, b_) do { *a = (uint64_t)(a_); *b = (uint64_t)(b_); return (code_); } while (0)

/* Reduced-erasure decode: copy present data, subtract present contributions
 * from the received repairs, invert the e x e Cauchy submatrix
 * B[t][s] = C[missing_s][repair_t], apply it (batch/codec.rs:278-406).
 * Check order: WrongOutputLen -> WrongCount -> per symbol IndexOutOfRange ->
 * DuplicateIndex -> WrongPayloadLen (batch/codec.rs:293-331). */
int or_gf256_decode(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                    const size_t *idx, const uint8_t *const *payloads,
                    const size_t *payload_lens, size_t count, uint8_t *out,
                    size_t out_len, uint64_t *a, uint64_t *b) {
    size_t n = k + m;
    if (out_len != k * L) RET(SCRS_DECODE_WRONG_OUTPUT_LEN, k * L, out_len);
    if (count != k) RET(SCRS_DECODE_WRONG_COUNT, k, count);
    unsigned char seen[256] = {0};
    size_t repair_cols[256], repair_src[256], e = 0;
    for (size_t s = 0; s < count; s++) {
        size_t i = idx[s];
        if (i >= n) RET(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i]) RET(SCRS_DECODE_DUPLICATE_INDEX, i, 0);
        seen[i] = 1;
        if (payload_lens[s] != L) RET(SCRS_DECODE_WRONG_PAYLOAD_LEN, L, payload_lens[s]);
        if (i < k) memcpy(out + i * L, payloads[s], L);
        else { repair_cols[e] = i - k; repair_src[e] = s; e++; }
    }
    if (e == 0) RET(SCRS_OK, 0, 0);
    size_t missing[256], nm = 0;
    for (size_t i = 0; i < k; i++) if (!seen[i]) missing[nm++] = i;
    uint8_t *work = (uint8_t *)malloc(e * L);
    for (size_t t = 0; t < e; t++) {
        memcpy(work + t * L, payloads[repair_src[t]], L);
        for (size_t s = 0; s < count; s++) {
            size_t i = idx[s];
            if (i >= k) continue;
            uint8_t c = coeffs[i * m + repair_cols[t]];
            for (size_t p = 0; p < L; p++) work[t * L + p] ^= or_gf256_mul(c, payloads[s][p]);
        }
    }
    uint8_t *bm = (uint8_t *)malloc(e * e), *bi = (uint8_t *)malloc(e * e);
    for (size_t t = 0; t < e; t++)
        for (size_t s = 0; s < e; s++) bm[t * e + s] = coeffs[missing[s] * m + repair_cols[t]];
    if (!gf256_invert(bm, e, bi)) {
        free(work); free(bm); free(bi);
        RET(SCRS_DECODE_INSUFFICIENT_RANK, e - 1, e);
    }
    for (size_t s = 0; s < e; s++) {
        uint8_t *dst = out + missing[s] * L;
        memset(dst, 0, L);
        for (size_t t = 0; t < e; t++) {
            uint8_t c = bi[s * e + t];
            for (size_t p = 0; p < L; p++) dst[p] ^= or_gf256_mul(c, work[t * L + p]);
        }
    }
    free(work); free(bm); free(bi);
    RET(SCRS_OK, 0, 0);
}

/* Full RREF of [M | P] with M[row][col] = G[col][idx] for G = [I | C]
 * (batch/codec.rs:412-461, decode_reference).  Returns the rank. */
int or_gf256_decode_rref(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                         const size_t *idx, const uint8_t *const *payloads,
                         uint8_t *out) {
    size_t stride = k + L;
    uint8_t *buf = (uint8_t *)calloc(k * stride, 1);
    for (size_t r = 0; r < k; r++) {
        if (idx[r] < k) buf[r * stride + idx[r]] = 1;
        else
            for (size_t c = 0; c < k; c++) buf[r * stride + c] = coeffs[c * m + (idx[r] - k)];
        memcpy(buf + r * stride + k, payloads[r], L);
    }
    size_t rank = 0;
    for (size_t col = 0; col < k && rank < k; col++) {
        size_t piv = rank;
        while (piv < k && buf[piv * stride + col] == 0) piv++;
        if (piv == k) continue;
        if (piv != rank)
            for (size_t c = 0; c < stride; c++) {
                uint8_t t = buf[rank * stride + c];
                buf[rank * stride + c] = buf[piv * stride + c];
                buf[piv * stride + c] = t;
            }
        uint8_t s = or_gf256_inv(buf[rank * stride + col]);
        for (size_t c = 0; c < stride; c++) buf[rank * stride + c] = or_gf256_mul(buf[rank * stride + c], s);
        for (size_t r = 0; r < k; r++) {
            if (r == rank) continue;
            uint8_t f = buf[r * stride + col];
            if (!f) continue;
            for (size_t c = 0; c < stride; c++) buf[r * stride + c] ^= or_gf256_mul(f, buf[rank * stride + c]);
        }
        rank++;
    }
    if (rank == k)
        for (size_t r = 0; r < k; r++) memcpy(out + r * L, buf + r * stride + k, L);
    free(buf);
    return (int)rank;
}

/* ====================================================================== */
/* GF(65536) Tower codec                                                   */
/* ====================================================================== */

static inline uint16_t ld16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static inline void st16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }

/* Feeding every data symbol i adds C[i][j] * data_i to repair j, elements
 * little-endian u16 (tower/encoder.rs:89-116; algebra/gf65536.rs:53-60). */
void or_tower_encode(const uint16_t *coeffs, size_t k, size_t m, size_t L,
                     const uint8_t *data, uint8_t *repairs) {
    memset(repairs, 0, m * L);
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++) {
            uint16_t c = coeffs[i * m + j];
            for (size_t p = 0; p + 1 < L; p += 2) {
                uint16_t v = ld16(repairs + j * L + p) ^ or_gf16_mul(c, ld16(data + i * L + p));
                st16(repairs + j * L + p, v);
            }
        }
}

/* Unique solution of the k received symbols by full Gauss-Jordan over
 * GF(2^16); the reference's rational-Lagrange decode (tower/decoder.rs:
 * 218-330) must agree byte for byte because the solution is unique. */
int or_tower_decode_rref(const uint16_t *coeffs, size_t k, size_t m, size_t L,
                         const size_t *idx, const uint8_t *const *payloads,
                         uint8_t *out) {
    size_t w = L / 2, stride = k + w;
    uint16_t *buf = (uint16_t *)calloc(k * stride, sizeof(uint16_t));
    for (size_t r = 0; r < k; r++) {
        if (idx[r] < k) buf[r * stride + idx[r]] = 1;
        else
            for (size_t c = 0; c < k; c++) buf[r * stride + c] = coeffs[c * m + (idx[r] - k)];
        for (size_t p = 0; p < w; p++) buf[r * stride + k + p] = ld16(payloads[r] + 2 * p);
    }
    size_t rank = 0;
    for (size_t col = 0; col < k && rank < k; col++) {
        size_t piv = rank;
        while (piv < k && buf[piv * stride + col] == 0) piv++;
        if (piv == k) continue;
        if (piv != rank)
            for (size_t c = 0; c < stride; c++) {
                uint16_t t = buf[rank * stride + c];
                buf[rank * stride + c] = buf[piv * stride + c];
                buf[piv * stride + c] = t;
            }
        uint16_t s = or_gf16_inv(buf[rank * stride + col]);
        for (size_t c = 0; c < stride; c++) buf[rank * stride + c] = or_gf16_mul(buf[rank * stride + c], s);
        for (size_t r = 0; r < k; r++) {
            if (r == rank) continue;
            uint16_t f = buf[r * stride + col];
            if (!f) continue;
            for (size_t c = 0; c < stride; c++) buf[r * stride + c] ^= or_gf16_mul(f, buf[rank * stride + c]);
        }
        rank++;
    }
    if (rank == k)
        for (size_t r = 0; r < k; r++)
            for (size_t p = 0; p < w; p++) st16(out + r * L + 2 * p, buf[r * stride + k + p]);
    free(buf);
    return (int)rank;
}

/* ====================================================================== */
/* GF(65536) additive-FFT codec: Reed-Solomon at points = wire index       */
/* ====================================================================== */

static uint16_t g16_exp[2 * 65535];
static uint32_t g16_log[65536];
static pthread_once_t g16_once = PTHREAD_ONCE_INIT;

/* log/exp over the generator 0x0108 (algebra/gf65536.rs:26), whose order
 * 65535 the KAT suite checks. */
static void g16_tables(void) {
    uint16_t v = 1;
    for (uint32_t i = 0; i < 65535; i++) {
        g16_exp[i] = g16_exp[i + 65535] = v;
        g16_log[v] = i;
        v = or_gf16_mul(v, 0x0108);
    }
}

static inline uint16_t g16_mul(uint16_t a, uint16_t b) {
    return (a && b) ? g16_exp[g16_log[a] + g16_log[b]] : 0;
}

static inline uint16_t g16_div(uint16_t a, uint16_t b) {
    return a ? g16_exp[g16_log[a] + 65535 - g16_log[b]] : 0;
}

/* Lagrange: P(x) = sum_t v_t A(x) / ((x + p_t) W_t), A(x) = prod_s (x + p_s),
 * W_t = prod_{s != t} (p_t + p_s).  When the points are exactly 0..k-1 with
 * k a power of two they form an additive subgroup, p_t + p_s runs over its
 * nonzero elements, and every W_t is their product. */
int or_rs16_interp(size_t k, size_t L, const size_t *pts, const uint8_t *const *vals,
                   size_t nout, const size_t *outs, uint8_t *out) {
    pthread_once(&g16_once, g16_tables);
    uint16_t *w = (uint16_t *)malloc(k * sizeof(uint16_t));
    uint16_t *coef = (uint16_t *)malloc(k * sizeof(uint16_t));
    int subgroup = (k & (k - 1)) == 0;
    for (size_t t = 0; t < k && subgroup; t++) subgroup = pts[t] == t;
    if (subgroup) {
        uint16_t prod = 1;
        for (size_t u = 1; u < k; u++) prod = g16_mul(prod, (uint16_t)u);
        for (size_t t = 0; t < k; t++) w[t] = prod;
    } else {
        for (size_t t = 0; t < k; t++) {
            uint16_t prod = 1;
            for (size_t s = 0; s < k; s++)
                if (s != t) prod = g16_mul(prod, (uint16_t)(pts[t] ^ pts[s]));
            if (!prod) {
                free(w);
                free(coef);
                return 0;
            }
            w[t] = prod;
        }
    }
    for (size_t i = 0; i < nout; i++) {
        const size_t x = outs[i];
        uint8_t *o = out + i * L;
        size_t hit = k;
        uint16_t a = 1;
        for (size_t s = 0; s < k; s++) {
            if (pts[s] == x) hit = s;
            else a = g16_mul(a, (uint16_t)(x ^ pts[s]));
        }
        if (hit < k) {
            memcpy(o, vals[hit], L);
            continue;
        }
        for (size_t t = 0; t < k; t++) coef[t] = g16_div(a, g16_mul((uint16_t)(x ^ pts[t]), w[t]));
        memset(o, 0, L);
        for (size_t t = 0; t < k; t++) {
            const uint32_t lc = g16_log[coef[t]];
            const uint8_t *v = vals[t];
            for (size_t p = 0; p + 1 < L; p += 2) {
                const uint16_t e = ld16(v + p);
                if (e) st16(o + p, (uint16_t)(ld16(o + p) ^ g16_exp[lc + g16_log[e]]));
            }
        }
    }
    free(w);
    free(coef);
    return 1;
}

void or_afft_encode(size_t k, size_t m, size_t L, const uint8_t *data, uint8_t *repairs) {
    size_t *pts = (size_t *)malloc(k * sizeof(size_t));
    size_t *outs = (size_t *)malloc(m * sizeof(size_t));
    const uint8_t **vals = (const uint8_t **)malloc(k * sizeof(uint8_t *));
    for (size_t i = 0; i < k; i++) {
        pts[i] = i;
        vals[i] = data + i * L;
    }
    for (size_t j = 0; j < m; j++) outs[j] = k + j;
    or_rs16_interp(k, L, pts, vals, m, outs, repairs);
    free(pts);
    free(outs);
    free(vals);
}

/* ====================================================================== */
/* CPU baseline: register-blocked source-major multiply-XOR                */
/* ====================================================================== */

/* Backend order GFNI+AVX2, then AVX2 nibble, then scalar
 * (simd/dispatch.rs:32-54). */
int or_cpu_plan(void) {
#if defined(__x86_64__)
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("gfni")) return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
#endif
    return 0;
}

static uint8_t MULTAB[256][256];
static int multab_ready = 0;
static pthread_once_t multab_once = PTHREAD_ONCE_INIT;
static void build_multab(void) {
    for (int a = 0; a < 256; a++)
        for (int b = 0; b < 256; b++) MULTAB[a][b] = or_gf256_mul((uint8_t)a, (uint8_t)b);
    multab_ready = 1;
}

#if defined(__x86_64__)
/* Four rows per 64-byte tile, every term folded into the eight tile
 * accumulators before one store (simd/x86.rs:658-760,
 * xor_scaled_bytes_rows_terms_gfni). */
__attribute__((target("avx2,gfni"))) static void rows_terms_gfni(
    uint8_t *dst, size_t L, size_t rows, size_t terms, const uint8_t *coef,
    const uint8_t *const *srcs) {
    size_t g = 0;
    for (; g + 4 <= rows; g += 4) {
        uint8_t *r0 = dst + g * L, *r1 = r0 + L, *r2 = r1 + L, *r3 = r2 + L;
        size_t p = 0;
        for (; p + 64 <= L; p += 64) {
            __m256i a00 = _mm256_loadu_si256((const __m256i *)(r0 + p));
            __m256i a01 = _mm256_loadu_si256((const __m256i *)(r0 + p + 32));
            __m256i a10 = _mm256_loadu_si256((const __m256i *)(r1 + p));
            __m256i a11 = _mm256_loadu_si256((const __m256i *)(r1 + p + 32));
            __m256i a20 = _mm256_loadu_si256((const __m256i *)(r2 + p));
            __m256i a21 = _mm256_loadu_si256((const __m256i *)(r2 + p + 32));
            __m256i a30 = _mm256_loadu_si256((const __m256i *)(r3 + p));
            __m256i a31 = _mm256_loadu_si256((const __m256i *)(r3 + p + 32));
            for (size_t t = 0; t < terms; t++) {
                const uint8_t *sp = srcs[t] + p;
                const uint8_t *cp = coef + t * rows + g;
                __m256i x0 = _mm256_loadu_si256((const __m256i *)sp);
                __m256i x1 = _mm256_loadu_si256((const __m256i *)(sp + 32));
                __m256i f0 = _mm256_set1_epi8((char)cp[0]), f1 = _mm256_set1_epi8((char)cp[1]);
                __m256i f2 = _mm256_set1_epi8((char)cp[2]), f3 = _mm256_set1_epi8((char)cp[3]);
                a00 = _mm256_xor_si256(a00, _mm256_gf2p8mul_epi8(x0, f0));
                a01 = _mm256_xor_si256(a01, _mm256_gf2p8mul_epi8(x1, f0));
                a10 = _mm256_xor_si256(a10, _mm256_gf2p8mul_epi8(x0, f1));
                a11 = _mm256_xor_si256(a11, _mm256_gf2p8mul_epi8(x1, f1));
                a20 = _mm256_xor_si256(a20, _mm256_gf2p8mul_epi8(x0, f2));
                a21 = _mm256_xor_si256(a21, _mm256_gf2p8mul_epi8(x1, f2));
                a30 = _mm256_xor_si256(a30, _mm256_gf2p8mul_epi8(x0, f3));
                a31 = _mm256_xor_si256(a31, _mm256_gf2p8mul_epi8(x1, f3));
            }
            _mm256_storeu_si256((__m256i *)(r0 + p), a00);
            _mm256_storeu_si256((__m256i *)(r0 + p + 32), a01);
            _mm256_storeu_si256((__m256i *)(r1 + p), a10);
            _mm256_storeu_si256((__m256i *)(r1 + p + 32), a11);
            _mm256_storeu_si256((__m256i *)(r2 + p), a20);
            _mm256_storeu_si256((__m256i *)(r2 + p + 32), a21);
            _mm256_storeu_si256((__m256i *)(r3 + p), a30);
            _mm256_storeu_si256((__m256i *)(r3 + p + 32), a31);
        }
        for (; p < L; p++)
            for (size_t t = 0; t < terms; t++)
                for (int q = 0; q < 4; q++)
                    dst[(g + q) * L + p] ^= MULTAB[coef[t * rows + g + q]][srcs[t][p]];
    }
    for (; g < rows; g++) {
        uint8_t *r0 = dst + g * L;
        size_t p = 0;
        for (; p + 64 <= L; p += 64) {
            __m256i a0 = _mm256_loadu_si256((const __m256i *)(r0 + p));
            __m256i a1 = _mm256_loadu_si256((const __m256i *)(r0 + p + 32));
            for (size_t t = 0; t < terms; t++) {
                __m256i f = _mm256_set1_epi8((char)coef[t * rows + g]);
                a0 = _mm256_xor_si256(a0, _mm256_gf2p8mul_epi8(_mm256_loadu_si256((const __m256i *)(srcs[t] + p)), f));
                a1 = _mm256_xor_si256(a1, _mm256_gf2p8mul_epi8(_mm256_loadu_si256((const __m256i *)(srcs[t] + p + 32)), f));
            }
            _mm256_storeu_si256((__m256i *)(r0 + p), a0);
            _mm256_storeu_si256((__m256i *)(r0 + p + 32), a1);
        }
        for (; p < L; p++)
            for (size_t t = 0; t < terms; t++) r0[p] ^= MULTAB[coef[t * rows + g]][srcs[t][p]];
    }
}

/* Split-nibble pshufb multiply (simd/x86.rs:1021-1060 and the ScaleTable of
 * simd/scale_table.rs:5-30), one row at a time. */
__attribute__((target("avx2"))) static void rows_terms_avx2(
    uint8_t *dst, size_t L, size_t rows, size_t terms, const uint8_t *coef,
    const uint8_t *const *srcs) {
    const __m256i mask = _mm256_set1_epi8(0x0F);
    for (size_t t = 0; t < terms; t++)
        for (size_t r = 0; r < rows; r++) {
            uint8_t c = coef[t * rows + r];
            uint8_t lo[16], hi[16];
            for (int i = 0; i < 16; i++) { lo[i] = MULTAB[c][i]; hi[i] = MULTAB[c][i << 4]; }
            __m256i tl = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)lo));
            __m256i th = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)hi));
            uint8_t *d = dst + r * L;
            size_t p = 0;
            for (; p + 32 <= L; p += 32) {
                __m256i x = _mm256_loadu_si256((const __m256i *)(srcs[t] + p));
                __m256i l = _mm256_shuffle_epi8(tl, _mm256_and_si256(x, mask));
                __m256i h = _mm256_shuffle_epi8(th, _mm256_and_si256(_mm256_srli_epi16(x, 4), mask));
                __m256i y = _mm256_xor_si256(_mm256_loadu_si256((const __m256i *)(d + p)), _mm256_xor_si256(l, h));
                _mm256_storeu_si256((__m256i *)(d + p), y);
            }
            for (; p < L; p++) d[p] ^= MULTAB[c][srcs[t][p]];
        }
}
#endif

void or_gf256_rows_terms_fast(uint8_t *dst, size_t L, size_t rows, size_t terms,
                              const uint8_t *coef, const uint8_t *const *srcs) {
    pthread_once(&multab_once, build_multab);
    int plan = or_cpu_plan();
#if defined(__x86_64__)
    if (plan == 2) { rows_terms_gfni(dst, L, rows, terms, coef, srcs); return; }
    if (plan == 1) { rows_terms_avx2(dst, L, rows, terms, coef, srcs); return; }
#endif
    (void)plan;
    for (size_t t = 0; t < terms; t++)
        for (size_t r = 0; r < rows; r++)
            for (size_t p = 0; p < L; p++) dst[r * L + p] ^= MULTAB[coef[t * rows + r]][srcs[t][p]];
}

/* Source-major encode of one codeword: zero the repairs, then one batched
 * rows_terms pass over all k sources (batch/codec.rs:156-193). */
static void encode_one(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                       const uint8_t *data, uint8_t *repairs) {
    const uint8_t *srcs[256];
    for (size_t i = 0; i < k; i++) srcs[i] = data + i * L;
    memset(repairs, 0, m * L);
    or_gf256_rows_terms_fast(repairs, L, m, k, coeffs, srcs);
}

typedef struct {
    const uint8_t *coeffs; size_t k, m, L; const uint8_t *data; uint8_t *repairs;
    size_t first, last;
} enc_job;

static void *enc_worker(void *p) {
    enc_job *j = (enc_job *)p;
    for (size_t c = j->first; c < j->last; c++)
        encode_one(j->coeffs, j->k, j->m, j->L, j->data + c * j->k * j->L, j->repairs + c * j->m * j->L);
    return NULL;
}

/* The reference's decode_into_with with the register-blocked kernel for
 * both payload passes: copy present data, subtract present contributions
 * from the received repairs in one rows_terms pass, invert, and apply the
 * e x e inverse (batch/codec.rs:278-406).  Inputs are already validated. */
static int decode_fast_one(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                           const uint16_t *idx, const uint8_t *const *payloads,
                           uint8_t *out, uint8_t *work) {
    unsigned char seen[256] = {0};
    size_t rc[256], e = 0, missing[256], nm = 0;
    const uint8_t *dsrc[256];
    size_t didx[256], nd = 0;
    for (size_t t = 0; t < k; t++) {
        size_t i = idx[t];
        seen[i] = 1;
        if (i < k) { memcpy(out + i * L, payloads[t], L); dsrc[nd] = payloads[t]; didx[nd++] = i; }
        else { memcpy(work + e * L, payloads[t], L); rc[e++] = i - k; }
    }
    if (e == 0) return 1;
    for (size_t i = 0; i < k; i++) if (!seen[i]) missing[nm++] = i;
    uint8_t flat[256 * 256];
    for (size_t w = 0; w < nd; w++)
        for (size_t t = 0; t < e; t++) flat[w * e + t] = coeffs[didx[w] * m + rc[t]];
    or_gf256_rows_terms_fast(work, L, e, nd, flat, dsrc);
    uint8_t bm[256 * 256], bi[256 * 256];
    for (size_t t = 0; t < e; t++)
        for (size_t s = 0; s < e; s++) bm[t * e + s] = coeffs[missing[s] * m + rc[t]];
    if (!gf256_invert(bm, e, bi)) return 0;
    /* out[missing_s] = sum_t Binv[s][t] work_t: terms are the e work rows. */
    const uint8_t *wsrc[256];
    uint8_t coef_tr[256 * 256];
    for (size_t t = 0; t < e; t++) {
        wsrc[t] = work + t * L;
        for (size_t s = 0; s < e; s++) coef_tr[t * e + s] = bi[s * e + t];
    }
    /* Rows are scattered in `out`; gather them through a contiguous block. */
    uint8_t *tmp = work + e * L;
    memset(tmp, 0, e * L);
    or_gf256_rows_terms_fast(tmp, L, e, e, coef_tr, wsrc);
    for (size_t s = 0; s < e; s++) memcpy(out + missing[s] * L, tmp + s * L, L);
    return 1;
}

typedef struct {
    const uint8_t *coeffs; size_t k, m, L; const uint16_t *idx;
    const uint8_t *data, *repairs; uint8_t *out; size_t first, last; int ok;
} dec_job;

static void *dec_worker(void *p) {
    dec_job *j = (dec_job *)p;
    size_t k = j->k, m = j->m, L = j->L;
    uint8_t *work = (uint8_t *)malloc(2 * (k < m ? k : m) * L + 1);
    const uint8_t *pl[256];
    j->ok = 1;
    for (size_t c = j->first; c < j->last; c++) {
        const uint16_t *idx = j->idx + c * k;
        for (size_t t = 0; t < k; t++)
            pl[t] = idx[t] < k ? j->data + (c * k + idx[t]) * L : j->repairs + (c * m + idx[t] - k) * L;
        if (!decode_fast_one(j->coeffs, k, m, L, idx, pl, j->out + c * k * L, work)) j->ok = 0;
    }
    free(work);
    return NULL;
}

int or_gf256_decode_many(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                         const uint16_t *idx, const uint8_t *data, const uint8_t *repairs,
                         uint8_t *out, size_t count, int threads) {
    pthread_once(&multab_once, build_multab);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > count) threads = (int)(count ? count : 1);
    pthread_t tid[256];
    dec_job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t] = (dec_job){coeffs, k, m, L, idx, data, repairs, out,
                            count * t / threads, count * (t + 1) / threads, 1};
        pthread_create(&tid[t], NULL, dec_worker, &jobs[t]);
    }
    int ok = 1;
    for (int t = 0; t < threads; t++) { pthread_join(tid[t], NULL); ok &= jobs[t].ok; }
    return ok;
}

void or_gf256_encode_many(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                          const uint8_t *data, uint8_t *repairs, size_t count,
                          int threads) {
    pthread_once(&multab_once, build_multab);
    if (threads < 1) threads = 1;
    if ((size_t)threads > count) threads = (int)(count ? count : 1);
    pthread_t tid[256];
    enc_job jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++) {
        jobs[t] = (enc_job){coeffs, k, m, L, data, repairs, count * t / threads, count * (t + 1) / threads};
        pthread_create(&tid[t], NULL, enc_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* ====================================================================== */
/* Seeded inputs                                                           */
/* ====================================================================== */

/* splitmix64 (Steele, Lea, Flood 2014). */
uint64_t or_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_fill_bytes(uint64_t seed, uint64_t byte_offset, uint8_t *out, size_t len) {
    for (size_t i = 0; i < len; i++) {
        uint64_t pos = byte_offset + i;
        uint64_t st = seed + (pos >> 3) * 0x9E3779B97F4A7C15ull;
        uint64_t w = or_splitmix64(&st);
        out[i] = (uint8_t)(w >> (8 * (pos & 7)));
    }
}
