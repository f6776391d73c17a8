/*
 * scrs_oracle.h - CPU restatement of the reference's payload arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library: only the test-suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  No product code path may call into it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  All arithmetic is
 * exact integer GF(2^n) math, so a correct implementation is bit-identical
 * to the reference on identical inputs.
 *
 * Pinning (what checks each function independently of its own formula):
 *   gf256_*          KATs of algebra/gf256.rs:428-465, generator cycle,
 *                    exhaustive 256x256 against the host's GF2P8MULB.
 *   gf16_*           KATs of algebra/gf65536.rs:221-236, primitive-element
 *                    order, Karatsuba vs schoolbook.
 *   *_cauchy_matrix  direct 1/(x+y) form vs the diagonal factorisation, and
 *                    the MDS property (every square minor nonsingular) on
 *                    small shapes.
 *   gf256_encode     closed form repair_j = sum_i C[i][j] data_i brute
 *                    forced per byte; the reference's own test vector
 *                    (batch/codec.rs:631-642).
 *   gf256_decode     decode(encode(x)) == x over every k-subset, and equality
 *                    with the independent full-RREF decode.
 *   tower_*          as above over GF(2^16).
 *   afft_*           Lagrange interpolation brute force on small shapes and
 *                    the reference's proptest regression seeds.
 */
#ifndef SCRS_ORACLE_H
#define SCRS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GF(2^8), polynomial 0x11B, generator 0x03 (algebra/gf256.rs) ---- */
uint8_t or_gf256_mul(uint8_t a, uint8_t b);
uint8_t or_gf256_inv(uint8_t a);
void or_gf256_exp_log(uint8_t exp_out[255], uint8_t log_out[256]);
/* Host GF2P8MULB product, or -1 when the CPU lacks GFNI. */
int or_gf256_mul_gfni(uint8_t a, uint8_t b);

/* ---- GF((2^8)^2), u^2 + u + 0x20 (algebra/gf65536.rs) ---- */
uint16_t or_gf16_mul(uint16_t a, uint16_t b);
uint16_t or_gf16_mul_schoolbook(uint16_t a, uint16_t b);
uint16_t or_gf16_square(uint16_t a);
uint16_t or_gf16_inv(uint16_t a);
uint16_t or_gf16_pow(uint16_t a, uint32_t e);

/* ---- Coefficient matrices, k x m source-major (i*m + j) ---- */
/* Returns 0 on success, or the engine capacity when k+m exceeds it. */
size_t or_std_cauchy_matrix(size_t k, size_t m, uint8_t *out);
size_t or_good_cauchy_matrix(size_t k, size_t m, uint8_t *out);
size_t or_good_cauchy_matrix_factored(size_t k, size_t m, uint8_t *out);
size_t or_tower_cauchy_matrix(size_t k, size_t m, uint16_t *out);

/* ---- GF(256) block codec (batch/codec.rs) ---- */
void or_gf256_encode(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                     const uint8_t *data, uint8_t *repairs);
/* Reduced-erasure decode with the reference's check order.  `idx` and
 * `payloads` hold `count` symbols; returns an scrs_code and fills *a, *b. */
int or_gf256_decode(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                    const size_t *idx, const uint8_t *const *payloads,
                    const size_t *payload_lens, size_t count, uint8_t *out,
                    size_t out_len, uint64_t *a, uint64_t *b);
/* Independent full Gauss-Jordan decode (batch/codec.rs:412-461). */
int or_gf256_decode_rref(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                         const size_t *idx, const uint8_t *const *payloads,
                         uint8_t *out);
/* dst[r] ^= sum_t coef[t*rows + r] * src[t], one scalar multiply at a time. */
void or_gf256_rows_terms(uint8_t *dst, size_t L, size_t rows, size_t terms,
                         const uint8_t *coef, const uint8_t *const *srcs);

/* ---- GF(65536) Tower codec (tower/encoder.rs, tower/decoder.rs) ---- */
void or_tower_encode(const uint16_t *coeffs, size_t k, size_t m, size_t L,
                     const uint8_t *data, uint8_t *repairs);
int or_tower_decode_rref(const uint16_t *coeffs, size_t k, size_t m, size_t L,
                         const size_t *idx, const uint8_t *const *payloads,
                         uint8_t *out);

/* ---- GF(65536) additive-FFT codec (afft/) ----
 * The AFFT engine is Reed-Solomon over GF(2^16) evaluated at the points
 * x = wire index (afft/profile.rs:41-46): data symbol i is P(i), repair j is
 * P(k + j), deg P < k.  Whatever transform computes it, the bytes are fixed
 * by that polynomial; the oracle evaluates it by Lagrange interpolation, the
 * closed form the reference's own encoder test uses (afft/encoder.rs:347-382). */

/* Evaluate the polynomial of degree < k through (pts[t], vals[t]), t < k,
 * at the `nout` points outs[i], writing L-byte symbols to out + i*L.
 * Returns 0 if the points are not distinct, else 1. */
int or_rs16_interp(size_t k, size_t L, const size_t *pts, const uint8_t *const *vals,
                   size_t nout, const size_t *outs, uint8_t *out);
/* Systematic encode: repairs[j] = P(k + j) with P(i) = data_i. */
void or_afft_encode(size_t k, size_t m, size_t L, const uint8_t *data, uint8_t *repairs);

/* ---- CPU baseline: the reference's x86 multiply-XOR plan ---- */
/* Returns the backend chosen: 2 = GFNI+AVX2, 1 = AVX2 nibble, 0 = scalar. */
int or_cpu_plan(void);
void or_gf256_rows_terms_fast(uint8_t *dst, size_t L, size_t rows, size_t terms,
                              const uint8_t *coef, const uint8_t *const *srcs);
/* Encode `count` codewords laid out back to back, on `threads` threads. */
void or_gf256_encode_many(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                          const uint8_t *data, uint8_t *repairs, size_t count,
                          int threads);

/* Decode `count` codewords laid out back to back (received symbol t of
 * codeword c is data row idx or repair row idx-k of that codeword), on
 * `threads` threads.  Returns 1 on success, 0 if a submatrix was singular. */
int or_gf256_decode_many(const uint8_t *coeffs, size_t k, size_t m, size_t L,
                         const uint16_t *idx, const uint8_t *data, const uint8_t *repairs,
                         uint8_t *out, size_t count, int threads);

/* ---- Seeded inputs shared with the benchmark ---- */
uint64_t or_splitmix64(uint64_t *state);
/* Byte i of the stream is byte (i % 8) of splitmix64 output word i / 8,
 * with the word counter starting at `seed` + word index (counter mode). */
void or_fill_bytes(uint64_t seed, uint64_t byte_offset, uint8_t *out, size_t len);

/* ---- CPU ports of the reference's GF(2^16) paths (scrs_cpu16.c) ---- */
int or_cpu16_plan(void); /* 2: AVX2+GFNI, 0: scalar */
void or_tower_encode_port(const uint16_t *coeffs, size_t k, size_t m, size_t L, const uint8_t *data,
                          uint8_t *repairs);
int or_tower_finalize_port(size_t k, size_t m, size_t L, const size_t *idx, const uint8_t *const *payloads,
                           uint8_t *out);
void or_tower_encode_port_many(const uint16_t *coeffs, size_t k, size_t m, size_t L, const uint8_t *data,
                               uint8_t *repairs, size_t count, int threads);
int or_tower_finalize_port_many(size_t k, size_t m, size_t L, const uint16_t *idx, const uint8_t *data,
                                const uint8_t *repairs, uint8_t *out, size_t count, int threads);
int or_afft_encode_port(size_t k, size_t m, size_t L, const uint8_t *data, uint8_t *repairs, int threads,
                        size_t max_strips);
void or_afft_novel_eval(const uint16_t *coef, size_t N, const uint16_t *points, size_t npoints, uint16_t *out);
int or_afft_decode_port(size_t k, size_t m, size_t L, const size_t *idx, const uint8_t *const *payloads,
                        uint8_t *out, int threads);

#ifdef __cplusplus
}
#endif

#endif /* SCRS_ORACLE_H */
