// kernels.hpp - launch interfaces of the GF(2^8) device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace scrs {

// Multiply-XOR over a batch of codewords:
//   dst[c][r] (op)= sum_t coef_c[t][r] * src[c][t]      for r < rows[c]
// Sources of codeword c are either strided (src_base + c * src_cw_stride +
// t * L) or read from src_ptrs[c * T + t].  Destinations likewise
// (dst_base + c * dst_cw_stride + r * L, or dst_ptrs[c * rmax + r]).  When
// copy_ptrs is set, copy_ptrs[c * T + t] (if non-null) receives a verbatim
// copy of source t: decode writes its present data rows this way while it
// streams them as terms.  Destinations are overwritten, never accumulated.
struct RowsTermsArgs {
    const uint8_t *src_base = nullptr;
    uint64_t src_cw_stride = 0;
    const uint8_t *const *src_ptrs = nullptr;
    uint8_t *dst_base = nullptr;
    uint64_t dst_cw_stride = 0;
    uint8_t *const *dst_ptrs = nullptr;
    uint8_t *const *copy_ptrs = nullptr;
    const uint32_t *rows = nullptr;   // per-codeword row count, or null for `rmax`
    const uint32_t *tables = nullptr; // product tables (build_nibble_tables layout)
    uint64_t table_cw_stride = 0;     // dwords between codewords' tables, 0 = shared
    uint32_t T = 0;                   // terms per codeword
    uint32_t rmax = 0;                // rows (upper bound when `rows` is set)
    uint32_t groups = 0;              // 16-row groups per term in the tables
    uint32_t groups_launched = 0;     // row groups this launch covers (set by the launcher)
    uint64_t L = 0;                   // symbol length in bytes
    uint64_t count = 0;               // codewords
};

// Table path: split-nibble product tables staged in LDS, 16-byte loads.
// `vec16` requires every row address and L to be multiples of 16.
hipError_t launch_rows_terms_lds(const RowsTermsArgs &a, bool vec16, hipStream_t s);

// Per-codeword decode solve.  For each codeword, from the k received
// indices: the missing data rows, the e x e inverse of the Cauchy
// submatrix, and the fused e x k coefficients applied directly to the k
// received symbols.  Writes packed nibble tables, row counts, and the
// source / destination / copy pointer tables consumed by the multiply-XOR.
struct DecodeSolveArgs {
    const uint16_t *indices = nullptr;   // [count][k], received order
    const uint8_t *coeffs = nullptr;     // k x m source-major
    const uint8_t *const *payload_ptrs = nullptr; // [count][k] or null: layout mode
    const uint8_t *data_base = nullptr;  // layout mode: data rows c*k*L + i*L
    const uint8_t *repair_base = nullptr;// layout mode: repair rows c*m*L + j*L
    uint8_t *out_base = nullptr;         // out rows c*k*L + i*L
    uint32_t *tables = nullptr;          // [count][k][groups][32][4]
    uint32_t *rows = nullptr;            // [count]
    const uint8_t **src_ptrs = nullptr;  // [count][k]
    uint8_t **dst_ptrs = nullptr;        // [count][rmax]
    uint8_t **copy_ptrs = nullptr;       // [count][k]
    uint32_t *error_flag = nullptr;      // set non-zero if a submatrix is singular
    uint32_t k = 0, m = 0, rmax = 0, groups = 0;
    uint64_t L = 0, count = 0;
};

hipError_t launch_decode_solve_gf256(const DecodeSolveArgs &a, hipStream_t s);

// Fill `len` bytes with the seeded splitmix64 counter stream: byte i is byte
// (i % 8) of splitmix64 word i / 8 drawn from state seed + (i/8) * golden.
hipError_t launch_fill_splitmix(uint8_t *dst, uint64_t len, uint64_t seed, uint64_t byte_offset,
                                hipStream_t s);

}  // namespace scrs
