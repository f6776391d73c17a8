/*
 * scrs_cpu16.c - CPU ports of the reference's GF(2^16) payload paths.
 *
 * TEST INFRASTRUCTURE ONLY: timed as the CPU baseline of the Tower and
 * additive-FFT workloads, and checked against the scalar oracle in
 * scrs_oracle.c.  Product code never links or calls it.
 *
 * Followed here (reference file:line, relative to the repository root):
 *   - GF(2^16) payload multiply-XOR on interleaved u16 with the AVX2+GFNI
 *     byte-swap form and the scalar Karatsuba fallback
 *     (scrs/src/tower/payload.rs:7-55, 157-239, 250-262, 399-477);
 *   - Tower incremental encode, one source over all repair rows
 *     (scrs/src/tower/encoder.rs:89-116) and the rational-Lagrange recipe and
 *     output-major apply of finalize (scrs/src/tower/decoder.rs:218-330,
 *     433-546);
 *   - the LCH additive FFT: subspace polynomials, heap-ordered node factors,
 *     node recursions and derivative (scrs/src/afft/transform.rs:41-60,
 *     188-212, 246-298, 301-559), the strip-blocked fused encode
 *     (scrs/src/afft/encoder.rs:29-36, 121-214) and the complete erasure-
 *     locator decode (scrs/src/afft/decoder.rs:269-356, 672-767).
 */
#include "scrs_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ---------------------------------------------------------------------- */
/* Field arithmetic as the reference's default build computes it           */
/* ---------------------------------------------------------------------- */

static uint8_t E8[255], L8[256];
static uint16_t E16[65535], L16[65536];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

static void tables_init(void) {
    or_gf256_exp_log(E8, L8);
    uint16_t v = 1;
    for (uint32_t i = 0; i < 65535u; i++) {
        E16[i] = v;
        L16[v] = (uint16_t)i;
        v = or_gf16_mul(v, 0x0108);
    }
}

/* GF(2^8) multiply through LOG/EXP with the zero short-circuit
 * (algebra/gf256.rs:135-146, the default `gf256-tables` feature). */
static inline uint8_t m8(uint8_t a, uint8_t b) {
    if (!a || !b) return 0;
    return E8[((unsigned)L8[a] + L8[b]) % 255u];
}

/* Karatsuba over GF((2^8)^2), u^2 = u + delta, delta = 0x20
 * (algebra/gf65536.rs:20, 79-90). */
static inline uint16_t m16(uint16_t x, uint16_t y) {
    uint8_t a = (uint8_t)x, b = (uint8_t)(x >> 8), c = (uint8_t)y, d = (uint8_t)(y >> 8);
    uint8_t ac = m8(a, c), bd = m8(b, d);
    uint8_t lo = ac ^ m8(0x20, bd);
    uint8_t hi = m8(a ^ b, c ^ d) ^ ac;
    return (uint16_t)(lo | (hi << 8));
}

static inline uint16_t ld16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static inline void st16(uint8_t *p, uint16_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}

static int have_gfni(void) {
#if defined(__x86_64__)
    static int v = -1;
    if (v < 0) v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("gfni");
    return v;
#else
    return 0;
#endif
}

int or_cpu16_plan(void) { return have_gfni() ? 2 : 0; }

/* ---------------------------------------------------------------------- */
/* Payload kernels (tower/payload.rs)                                      */
/* ---------------------------------------------------------------------- */

static void axpy_scalar(uint8_t *dst, uint16_t c, const uint8_t *src, size_t len) {
    if (c == 0) return;
    if (c == 1) {
        for (size_t i = 0; i < len; i++) dst[i] ^= src[i];
        return;
    }
    for (size_t i = 0; i + 1 < len; i += 2) st16(dst + i, ld16(dst + i) ^ m16(ld16(src + i), c));
}

static void fwd_scalar(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    for (size_t i = 0; i + 1 < len; i += 2) {
        uint16_t l = ld16(lo + i), h = ld16(hi + i);
        uint16_t nl = l ^ m16(c, h);
        st16(lo + i, nl);
        st16(hi + i, h ^ nl);
    }
}

static void inv_scalar(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    for (size_t i = 0; i + 1 < len; i += 2) {
        uint16_t l = ld16(lo + i), h = ld16(hi + i);
        uint16_t nh = h ^ l;
        st16(lo + i, l ^ m16(c, nh));
        st16(hi + i, nh);
    }
}

static void xor_into(uint8_t *dst, const uint8_t *src, size_t len) {
    for (size_t i = 0; i < len; i++) dst[i] ^= src[i];
}

#if defined(__x86_64__)
/* Multiplication of interleaved [a, b] by c + d u is
 * [c a + delta d b, d a + (c + d) b]: the source times [c, c+d] plus the
 * adjacent-byte-swapped source times [delta d, d], bytewise (payload.rs:250-262,
 * 399-411). */
static void factor_words(uint16_t coef, int16_t *same, int16_t *cross) {
    uint8_t c0 = (uint8_t)coef, c1 = (uint8_t)(coef >> 8);
    *same = (int16_t)(uint16_t)(c0 | ((c0 ^ c1) << 8));
    *cross = (int16_t)(uint16_t)(m8(0x20, c1) | (c1 << 8));
}

__attribute__((target("avx2,gfni"))) static inline __m256i scaled_vec(__m256i s, __m256i same, __m256i cross) {
    const __m256i swap = _mm256_setr_epi8(1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 1, 0, 3, 2, 5, 4, 7,
                                          6, 9, 8, 11, 10, 13, 12, 15, 14);
    __m256i sw = _mm256_shuffle_epi8(s, swap);
    return _mm256_xor_si256(_mm256_gf2p8mul_epi8(s, same), _mm256_gf2p8mul_epi8(sw, cross));
}

__attribute__((target("avx2,gfni"))) static void axpy_gfni(uint8_t *dst, uint16_t c, const uint8_t *src,
                                                          size_t len) {
    int16_t a, b;
    factor_words(c, &a, &b);
    const __m256i same = _mm256_set1_epi16(a), cross = _mm256_set1_epi16(b);
    size_t v = len / 32 * 32;
    for (size_t o = 0; o < v; o += 32) {
        __m256i s = _mm256_loadu_si256((const __m256i *)(src + o));
        __m256i d = _mm256_loadu_si256((const __m256i *)(dst + o));
        _mm256_storeu_si256((__m256i *)(dst + o), _mm256_xor_si256(d, scaled_vec(s, same, cross)));
    }
    axpy_scalar(dst + v, c, src + v, len - v);
}

/* xor_scaled_bytes_rows_gfni (payload.rs:433-477): each 32-byte source
 * vector is loaded once and folded into every row. */
__attribute__((target("avx2,gfni"))) static void rows_gfni(uint8_t *dsts, size_t L, const uint16_t *coefs,
                                                          size_t rows, const uint8_t *src) {
    size_t v = L / 32 * 32;
    for (size_t o = 0; o < v; o += 32) {
        __m256i s = _mm256_loadu_si256((const __m256i *)(src + o));
        for (size_t r = 0; r < rows; r++) {
            uint16_t c = coefs[r];
            if (c == 0) continue;
            uint8_t *dp = dsts + r * L + o;
            __m256i d = _mm256_loadu_si256((const __m256i *)dp);
            __m256i x;
            if (c == 1) {
                x = s;
            } else {
                int16_t a, b;
                factor_words(c, &a, &b);
                x = scaled_vec(s, _mm256_set1_epi16(a), _mm256_set1_epi16(b));
            }
            _mm256_storeu_si256((__m256i *)dp, _mm256_xor_si256(d, x));
        }
    }
    if (v != L)
        for (size_t r = 0; r < rows; r++) axpy_scalar(dsts + r * L + v, coefs[r], src + v, L - v);
}

__attribute__((target("avx2,gfni"))) static void fwd_gfni(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    int16_t a, b;
    factor_words(c, &a, &b);
    const __m256i same = _mm256_set1_epi16(a), cross = _mm256_set1_epi16(b);
    size_t v = len / 32 * 32;
    for (size_t o = 0; o < v; o += 32) {
        __m256i l = _mm256_loadu_si256((const __m256i *)(lo + o));
        __m256i h = _mm256_loadu_si256((const __m256i *)(hi + o));
        __m256i nl = _mm256_xor_si256(l, scaled_vec(h, same, cross));
        _mm256_storeu_si256((__m256i *)(lo + o), nl);
        _mm256_storeu_si256((__m256i *)(hi + o), _mm256_xor_si256(h, nl));
    }
    fwd_scalar(lo + v, hi + v, c, len - v);
}

__attribute__((target("avx2,gfni"))) static void inv_gfni(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    int16_t a, b;
    factor_words(c, &a, &b);
    const __m256i same = _mm256_set1_epi16(a), cross = _mm256_set1_epi16(b);
    size_t v = len / 32 * 32;
    for (size_t o = 0; o < v; o += 32) {
        __m256i l = _mm256_loadu_si256((const __m256i *)(lo + o));
        __m256i h = _mm256_loadu_si256((const __m256i *)(hi + o));
        __m256i nh = _mm256_xor_si256(h, l);
        _mm256_storeu_si256((__m256i *)(lo + o), _mm256_xor_si256(l, scaled_vec(nh, same, cross)));
        _mm256_storeu_si256((__m256i *)(hi + o), nh);
    }
    inv_scalar(lo + v, hi + v, c, len - v);
}
#endif

/* xor_scaled_bytes (payload.rs:7-55). */
static void axpy16(uint8_t *dst, uint16_t c, const uint8_t *src, size_t len) {
    if (c == 0) return;
    if (c == 1) {
        xor_into(dst, src, len);
        return;
    }
#if defined(__x86_64__)
    if (have_gfni()) {
        axpy_gfni(dst, c, src, len);
        return;
    }
#endif
    axpy_scalar(dst, c, src, len);
}

/* xor_scaled_bytes_rows (payload.rs:157-185). */
static void rows16(uint8_t *dsts, size_t L, const uint16_t *coefs, size_t rows, const uint8_t *src) {
#if defined(__x86_64__)
    if (have_gfni() && L >= 32) {
        rows_gfni(dsts, L, coefs, rows, src);
        return;
    }
#endif
    for (size_t r = 0; r < rows; r++) axpy16(dsts + r * L, coefs[r], src, L);
}

/* fused_forward_bytes / fused_inverse_bytes (payload.rs:58-136): a zero
 * factor reduces both to high ^= low. */
static void fused_fwd(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    if (c == 0) {
        xor_into(hi, lo, len);
        return;
    }
#if defined(__x86_64__)
    if (have_gfni()) {
        fwd_gfni(lo, hi, c, len);
        return;
    }
#endif
    fwd_scalar(lo, hi, c, len);
}

static void fused_inv(uint8_t *lo, uint8_t *hi, uint16_t c, size_t len) {
    if (c == 0) {
        xor_into(hi, lo, len);
        return;
    }
#if defined(__x86_64__)
    if (have_gfni()) {
        inv_gfni(lo, hi, c, len);
        return;
    }
#endif
    inv_scalar(lo, hi, c, len);
}

/* ---------------------------------------------------------------------- */
/* Tower: incremental encode and finalize                                  */
/* ---------------------------------------------------------------------- */

/* Feeding all k data symbols in index order (tower/encoder.rs:89-116):
 * repairs start at zero and data symbol i is folded into every repair row
 * with the coefficient row i of the k x m Tower Cauchy matrix. */
void or_tower_encode_port(const uint16_t *coeffs, size_t k, size_t m, size_t L, const uint8_t *data,
                          uint8_t *repairs) {
    pthread_once(&tables_once, tables_init);
    memset(repairs, 0, m * L);
    for (size_t i = 0; i < k; i++) rows16(repairs, L, coeffs + i * m, m, data + i * L);
}

static inline uint16_t inv16(uint16_t a) { return a ? E16[(65535u - L16[a]) % 65535u] : 0; }
static inline uint16_t tpow(size_t e) { return E16[e % 65535u]; } /* g^e, g = 0x0108 */

/* finalize_into_with (tower/decoder.rs:200-330) on k received symbols
 * idx[t] with payloads[t].  The repairs used are the first r received in
 * index order; the recipe is the rational-Lagrange closed form
 * (tower/decoder.rs:433-546) over x_i = g^i and y_j = g^(k+j)
 * (tower/cauchy.rs:46-53); the apply is output-major.  Returns 0, or -1 when
 * the received set has too few repairs. */
int or_tower_finalize_port(size_t k, size_t m, size_t L, const size_t *idx, const uint8_t *const *payloads,
                           uint8_t *out) {
    pthread_once(&tables_once, tables_init);
    const size_t n = k + m;
    const uint8_t **at = (const uint8_t **)calloc(n, sizeof(*at));
    size_t *missing = (size_t *)malloc(k * sizeof(size_t)), *present = (size_t *)malloc(k * sizeof(size_t));
    size_t *reps = (size_t *)malloc((m + 1) * sizeof(size_t));
    for (size_t t = 0; t < k; t++) at[idx[t]] = payloads[t];
    size_t r = 0, np = 0, nr = 0;
    for (size_t i = 0; i < k; i++) {
        if (at[i]) present[np++] = i;
        else missing[r++] = i;
    }
    for (size_t j = 0; j < m && nr < r; j++)
        if (at[k + j]) reps[nr++] = j;
    int rc = 0;
    if (nr != r) {
        rc = -1;
        goto done;
    }
    for (size_t p = 0; p < np; p++) memcpy(out + present[p] * L, at[present[p]], L);
    if (r) {
        uint16_t *rv = (uint16_t *)malloc(r * sizeof(uint16_t)), *cv = (uint16_t *)malloc(r * sizeof(uint16_t));
        uint16_t *rf = (uint16_t *)malloc(r * sizeof(uint16_t)), *cf = (uint16_t *)malloc(r * sizeof(uint16_t));
        uint16_t *inverse = (uint16_t *)malloc(r * r * sizeof(uint16_t));
        uint16_t *pres = (uint16_t *)malloc((np ? np : 1) * r * sizeof(uint16_t));
        for (size_t i = 0; i < r; i++) rv[i] = tpow(k + reps[i]);
        for (size_t j = 0; j < r; j++) cv[j] = tpow(missing[j]);
        for (size_t i = 0; i < r; i++) {
            uint16_t cross = 1, within = 1;
            for (size_t j = 0; j < r; j++) cross = m16(cross, rv[i] ^ cv[j]);
            for (size_t l = 0; l < r; l++)
                if (l != i) within = m16(within, rv[i] ^ rv[l]);
            rf[i] = m16(cross, inv16(within));
        }
        for (size_t j = 0; j < r; j++) {
            uint16_t cross = 1, within = 1;
            for (size_t i = 0; i < r; i++) cross = m16(cross, cv[j] ^ rv[i]);
            for (size_t h = 0; h < r; h++)
                if (h != j) within = m16(within, cv[j] ^ cv[h]);
            cf[j] = m16(cross, inv16(within));
        }
        for (size_t j = 0; j < r; j++)
            for (size_t i = 0; i < r; i++) inverse[j * r + i] = m16(m16(rf[i], cf[j]), inv16(cv[j] ^ rv[i]));
        for (size_t p = 0; p < np; p++) {
            uint16_t z = tpow(present[p]), rows_prod = 1;
            for (size_t i = 0; i < r; i++) rows_prod = m16(rows_prod, z ^ rv[i]);
            uint16_t rinv = inv16(rows_prod);
            for (size_t j = 0; j < r; j++) {
                uint16_t others = 1;
                for (size_t h = 0; h < r; h++)
                    if (h != j) others = m16(others, z ^ cv[h]);
                pres[p * r + j] = m16(m16(cf[j], others), rinv);
            }
        }
        for (size_t j = 0; j < r; j++) {
            uint8_t *row = out + missing[j] * L;
            memset(row, 0, L);
            for (size_t i = 0; i < r; i++) axpy16(row, inverse[j * r + i], at[k + reps[i]], L);
            for (size_t p = 0; p < np; p++) axpy16(row, pres[p * r + j], at[present[p]], L);
        }
        free(rv);
        free(cv);
        free(rf);
        free(cf);
        free(inverse);
        free(pres);
    }
done:
    free(at);
    free(missing);
    free(present);
    free(reps);
    return rc;
}

typedef struct {
    const uint16_t *coeffs;
    size_t k, m, L, count;
    const uint16_t *idx; /* count x k, or null: encode */
    const uint8_t *data, *repairs;
    uint8_t *out;
    int t, threads, ok;
} tower_job;

static void *tower_worker(void *arg) {
    tower_job *j = (tower_job *)arg;
    size_t k = j->k, m = j->m, L = j->L;
    size_t *ix = (size_t *)malloc(k * sizeof(size_t));
    const uint8_t **pp = (const uint8_t **)malloc(k * sizeof(*pp));
    j->ok = 1;
    for (size_t c = (size_t)j->t; c < j->count; c += (size_t)j->threads) {
        if (!j->idx) {
            or_tower_encode_port(j->coeffs, k, m, L, j->data + c * k * L, j->out + c * m * L);
            continue;
        }
        for (size_t t = 0; t < k; t++) {
            size_t x = j->idx[c * k + t];
            ix[t] = x;
            pp[t] = x < k ? j->data + (c * k + x) * L : j->repairs + (c * m + (x - k)) * L;
        }
        if (or_tower_finalize_port(k, m, L, ix, pp, j->out + c * k * L) != 0) j->ok = 0;
    }
    free(ix);
    free(pp);
    return NULL;
}

static int tower_many(const uint16_t *coeffs, size_t k, size_t m, size_t L, const uint16_t *idx,
                      const uint8_t *data, const uint8_t *repairs, uint8_t *out, size_t count, int threads) {
    pthread_once(&tables_once, tables_init);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > count) threads = (int)(count ? count : 1);
    pthread_t tid[256];
    tower_job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t] = (tower_job){coeffs, k, m, L, count, idx, data, repairs, out, t, threads, 1};
        pthread_create(&tid[t], NULL, tower_worker, &jobs[t]);
    }
    int ok = 1;
    for (int t = 0; t < threads; t++) {
        pthread_join(tid[t], NULL);
        ok &= jobs[t].ok;
    }
    return ok;
}

/* `count` codewords back to back, one codeword per thread at a time. */
void or_tower_encode_port_many(const uint16_t *coeffs, size_t k, size_t m, size_t L, const uint8_t *data,
                               uint8_t *repairs, size_t count, int threads) {
    tower_many(coeffs, k, m, L, NULL, data, NULL, repairs, count, threads);
}

/* Codeword c receives idx[c*k .. c*k+k) (symbol i < k read from
 * data + (c*k+i)*L, repair j from repairs + (c*m+j)*L); 1 when all decoded. */
int or_tower_finalize_port_many(size_t k, size_t m, size_t L, const uint16_t *idx, const uint8_t *data,
                                const uint8_t *repairs, uint8_t *out, size_t count, int threads) {
    return tower_many(NULL, k, m, L, idx, data, repairs, out, count, threads);
}

/* ---------------------------------------------------------------------- */
/* Additive FFT                                                            */
/* ---------------------------------------------------------------------- */

static uint16_t sq16(uint16_t a) { return m16(a, a); }

/* Evaluate the linearized polynomial sum_j c_j x^(2^j). */
static uint16_t lin_eval(const uint16_t *c, size_t nc, uint16_t x) {
    uint16_t r = 0, p = x;
    for (size_t j = 0; j < nc; j++) {
        r ^= m16(c[j], p);
        p = sq16(p);
    }
    return r;
}

/* Subspace polynomials W_0(x) = x, W_{i+1} = W_i^2 + W_i(b_i) W_i with
 * b_i = 2^i, normalised by W_i(b_i) (afft/transform.rs:246-268).  coef
 * holds `levels` rows of 17 coefficients; ninv[i] = 1 / W_i(b_i). */
static void subspace_polys(size_t levels, uint16_t coef[][17], uint16_t *ninv) {
    uint16_t cur[17] = {1}, next[17];
    size_t nc = 1;
    for (size_t i = 0; i < levels; i++) {
        uint16_t norm = lin_eval(cur, nc, (uint16_t)(1u << i));
        memcpy(coef[i], cur, sizeof(cur));
        ninv[i] = inv16(norm);
        memset(next, 0, sizeof(next));
        for (size_t j = 0; j < nc; j++) {
            next[j] ^= m16(norm, cur[j]);
            next[j + 1] ^= sq16(cur[j]);
        }
        memcpy(cur, next, sizeof(cur));
        nc++;
    }
}

/* Heap-ordered node factors (fill_factors, afft/transform.rs:281-298) and
 * the derivative factors W_i'(0) / W_i(b_i) (:41-60). */
static void fill_factors(uint16_t *f, uint16_t coef[][17], const uint16_t *ninv, size_t node, size_t dim,
                         uint16_t shift) {
    f[node] = m16(lin_eval(coef[dim - 1], dim, shift), ninv[dim - 1]);
    if (dim > 1) {
        fill_factors(f, coef, ninv, 2 * node, dim - 1, shift);
        fill_factors(f, coef, ninv, 2 * node + 1, dim - 1, (uint16_t)(shift ^ (1u << (dim - 1))));
    }
}

static void afft_plan(size_t logn, uint16_t *factors, uint16_t *dfac) {
    uint16_t coef[16][17], ninv[16];
    subspace_polys(logn, coef, ninv);
    if (logn) fill_factors(factors, coef, ninv, 1, logn, 0);
    for (size_t i = 0; i < logn; i++) dfac[i] = m16(coef[i][0], ninv[i]);
}

/* Node recursions over rows of w bytes (afft/transform.rs:335-345, 347-386, 541-551). */
static void fwd_node(uint8_t *rows, size_t nrows, size_t w, const uint16_t *f, size_t node, size_t dim) {
    size_t half = nrows / 2;
    fused_fwd(rows, rows + half * w, f[node], half * w);
    if (dim > 1) {
        fwd_node(rows, half, w, f, 2 * node, dim - 1);
        fwd_node(rows + half * w, half, w, f, 2 * node + 1, dim - 1);
    }
}

static void fwd_range_node(uint8_t *rows, size_t nrows, size_t w, const uint16_t *f, size_t node, size_t dim,
                           size_t r0, size_t r1) {
    size_t half = nrows / 2;
    fused_fwd(rows, rows + half * w, f[node], half * w);
    if (dim <= 1) return;
    if (r0 < half) fwd_range_node(rows, half, w, f, 2 * node, dim - 1, r0, r1 < half ? r1 : half);
    if (r1 > half)
        fwd_range_node(rows + half * w, half, w, f, 2 * node + 1, dim - 1, r0 > half ? r0 - half : 0, r1 - half);
}

static void inv_node(uint8_t *rows, size_t nrows, size_t w, const uint16_t *f, size_t node, size_t dim) {
    size_t half = nrows / 2;
    if (dim > 1) {
        inv_node(rows, half, w, f, 2 * node, dim - 1);
        inv_node(rows + half * w, half, w, f, 2 * node + 1, dim - 1);
    }
    fused_inv(rows, rows + half * w, f[node], half * w);
}

/* derivative_bytes (afft/transform.rs:188-212). */
static void derivative(const uint8_t *coefs, size_t nrows, size_t w, const uint16_t *dfac, uint8_t *out) {
    memset(out, 0, nrows * w);
    for (size_t i = 1; i < nrows; i++)
        for (size_t rem = i; rem; rem &= rem - 1) {
            size_t bit = (size_t)__builtin_ctzl(rem);
            axpy16(out + (i ^ ((size_t)1 << bit)) * w, dfac[bit], coefs + i * w, w);
        }
}

static size_t log2z(size_t x) {
    size_t d = 0;
    while (((size_t)1 << d) < x) d++;
    return d;
}

/* strip_width (afft/encoder.rs:29-36): 768 KiB working set per strip. */
static size_t strip_width(size_t rows_per_strip, size_t L) {
    size_t cap = (786432 / rows_per_strip) & ~(size_t)1;
    if (cap < 2) cap = 2;
    return cap < L ? cap : L;
}

typedef struct {
    size_t k, m, L, w, strips, t, threads;
    const uint16_t *fi, *ft;
    const uint8_t *data;
    uint8_t *repairs;
} enc_strip_job;

static void *enc_strip_worker(void *arg) {
    enc_strip_job *j = (enc_strip_job *)arg;
    size_t k = j->k, m = j->m, L = j->L, w = j->w;
    uint8_t *strip = (uint8_t *)malloc(k * w);
    for (size_t s = j->t; s < j->strips; s += j->threads) {
        size_t col = s * w, ww = L - col < w ? L - col : w;
        for (size_t r = 0; r < k; r++) memcpy(strip + r * ww, j->data + r * L + col, ww);
        /* inverse_truncated_bytes with active == k == padded_k is the full
         * inverse (transform.rs:466-468), then the repair coset is the
         * forward subtree at node 3 (:142-165). */
        inv_node(strip, k, ww, j->fi, 1, log2z(k));
        fwd_range_node(strip, k, ww, j->ft, 3, log2z(2 * k) - 1, 0, m);
        for (size_t r = 0; r < m; r++) memcpy(j->repairs + r * L + col, strip + r * ww, ww);
    }
    free(strip);
    return NULL;
}

/* Fused strip-blocked encode (afft/encoder.rs:150-214) for k a power of two
 * and m <= k; columns are cut into the reference's strips and strips are
 * shared out over `threads`.  `max_strips` > 0 processes only the first
 * strips (a bounded sample).  Returns -1 for geometries off the fused path. */
int or_afft_encode_port(size_t k, size_t m, size_t L, const uint8_t *data, uint8_t *repairs, int threads,
                        size_t max_strips) {
    pthread_once(&tables_once, tables_init);
    if (k < 2 || (k & (k - 1)) || m > k || L % 2 || k * 2 > 65536) return -1;
    size_t lk = log2z(k);
    uint16_t *fi = (uint16_t *)calloc(k, sizeof(uint16_t)), *ft = (uint16_t *)calloc(2 * k, sizeof(uint16_t));
    uint16_t dfac[17];
    afft_plan(lk, fi, dfac);
    afft_plan(lk + 1, ft, dfac);
    size_t w = strip_width(k, L);
    size_t strips = (L + w - 1) / w;
    if (max_strips && max_strips < strips) strips = max_strips;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > strips) threads = (int)strips;
    pthread_t tid[256];
    enc_strip_job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t] = (enc_strip_job){k, m, L, w, strips, (size_t)t, (size_t)threads, fi, ft, data, repairs};
        pthread_create(&tid[t], NULL, enc_strip_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    free(fi);
    free(ft);
    return (int)w;
}

/* Walsh-Hadamard transform modulo 65535 (afft/decoder.rs:732-754). */
static void wht(uint32_t *v, size_t n) {
    for (size_t half = 1; half < n; half *= 2)
        for (size_t b = 0; b < n; b += 2 * half)
            for (size_t p = 0; p < half; p++) {
                uint32_t l = v[b + p], r = v[b + half + p], s = l + r;
                v[b + p] = s >= 65535u ? s - 65535u : s;
                v[b + half + p] = l >= r ? l - r : l + 65535u - r;
            }
}

/* locator_evaluations_into (afft/decoder.rs:672-713): for every point x,
 * log of the product of (x + e) over erased e != x, by two transforms of the
 * erased-set indicator and of the log table.  values[x] (known x) is that
 * product, derivs[x] (erased x) likewise. */
static void locator(const uint8_t *known, size_t n, uint16_t *values, uint16_t *derivs) {
    uint32_t *ind = (uint32_t *)malloc(n * 4), *lg = (uint32_t *)malloc(n * 4);
    for (size_t x = 0; x < n; x++) {
        ind[x] = known[x] ? 0 : 1;
        lg[x] = x ? L16[x] : 0;
    }
    wht(ind, n);
    wht(lg, n);
    for (size_t x = 0; x < n; x++) ind[x] = (uint32_t)(((uint64_t)ind[x] * lg[x]) % 65535u);
    wht(ind, n);
    uint64_t inv_n = 1, base = 32768;
    for (size_t e = log2z(n); e; e--) inv_n = inv_n * base % 65535u;
    for (size_t x = 0; x < n; x++) {
        uint16_t p = E16[(uint32_t)((ind[x] * inv_n) % 65535u)];
        values[x] = known[x] ? p : 0;
        derivs[x] = known[x] ? 0 : p;
    }
    free(ind);
    free(lg);
}

typedef struct {
    size_t k, n, N, L, c0, c1;
    const uint16_t *f, *dfac, *values, *derivs;
    const uint8_t *const *at;
    const uint8_t *known;
    uint8_t *out;
} dec_slice_job;

static void *dec_slice_worker(void *arg) {
    dec_slice_job *j = (dec_slice_job *)arg;
    size_t w = j->c1 - j->c0, N = j->N;
    if (!w) return NULL;
    uint8_t *w0 = (uint8_t *)calloc(N * w, 1), *w1 = (uint8_t *)malloc(N * w);
    for (size_t x = 0; x < j->n; x++)
        if (j->at[x]) axpy16(w0 + x * w, j->values[x], j->at[x] + j->c0, w);
    inv_node(w0, N, w, j->f, 1, log2z(N));
    derivative(w0, N, w, j->dfac, w1);
    fwd_node(w1, N, w, j->f, 1, log2z(N));
    for (size_t d = 0; d < j->k; d++) {
        uint8_t *o = j->out + d * j->L + j->c0;
        if (j->known[d]) {
            memcpy(o, j->at[d] + j->c0, w);
        } else {
            memset(o, 0, w);
            axpy16(o, inv16(j->derivs[d]), w1 + d * w, w);
        }
    }
    free(w0);
    free(w1);
    return NULL;
}

/* Complete decode (afft/decoder.rs:269-356) of k received symbols
 * (idx[t], payloads[t]): locator products scale the received evaluations,
 * then inverse transform, formal derivative, forward transform, and each
 * missing data row is divided by its locator derivative.  Column slices of
 * even width are shared out over `threads` (each runs the whole pipeline on
 * its slice).  Returns 0, or -1 for a bad geometry. */
int or_afft_decode_port(size_t k, size_t m, size_t L, const size_t *idx, const uint8_t *const *payloads,
                        uint8_t *out, int threads) {
    pthread_once(&tables_once, tables_init);
    size_t n = k + m, N = 1;
    while (N < n) N *= 2;
    if (n > 65536 || L % 2) return -1;
    uint8_t *known = (uint8_t *)calloc(N, 1);
    const uint8_t **at = (const uint8_t **)calloc(n, sizeof(*at));
    for (size_t t = 0; t < k; t++) {
        known[idx[t]] = 1;
        at[idx[t]] = payloads[t];
    }
    uint16_t *values = (uint16_t *)malloc(N * 2), *derivs = (uint16_t *)malloc(N * 2);
    uint16_t *f = (uint16_t *)calloc(N, 2), dfac[17];
    locator(known, N, values, derivs);
    afft_plan(log2z(N), f, dfac);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    size_t pairs = L / 2;
    if ((size_t)threads > pairs) threads = (int)(pairs ? pairs : 1);
    pthread_t tid[256];
    dec_slice_job jobs[256];
    for (int t = 0; t < threads; t++) {
        size_t c0 = 2 * (pairs * t / threads), c1 = 2 * (pairs * (t + 1) / threads);
        jobs[t] = (dec_slice_job){k, n, N, L, c0, c1, f, dfac, values, derivs, at, known, out};
        pthread_create(&tid[t], NULL, dec_slice_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    free(known);
    free(at);
    free(values);
    free(derivs);
    free(f);
    return 0;
}

/* Direct evaluation in the novel basis (the reference's check,
 * afft/transform.rs:565-590): out[p] = sum_i coef[i] X_i(points[p]) with
 * X_i(x) the product of W_b(x) / W_b(2^b) over the set bits b of i. */
void or_afft_novel_eval(const uint16_t *coef, size_t N, const uint16_t *points, size_t npoints, uint16_t *out) {
    pthread_once(&tables_once, tables_init);
    size_t logn = log2z(N);
    uint16_t c[16][17], ninv[16], wbar[16];
    subspace_polys(logn, c, ninv);
    uint16_t *X = (uint16_t *)malloc(N * sizeof(uint16_t));
    for (size_t p = 0; p < npoints; p++) {
        for (size_t b = 0; b < logn; b++) wbar[b] = m16(lin_eval(c[b], b + 1, points[p]), ninv[b]);
        X[0] = 1;
        uint16_t r = m16(coef[0], 1);
        for (size_t i = 1; i < N; i++) {
            X[i] = m16(X[i & (i - 1)], wbar[__builtin_ctzl(i)]);
            r ^= m16(coef[i], X[i]);
        }
        out[p] = r;
    }
    free(X);
}
