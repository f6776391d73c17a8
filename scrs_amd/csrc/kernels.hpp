// kernels.hpp - launch interfaces of the device kernels.
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
// streams them as terms.  Destinations are overwritten unless `accumulate`
// is set, in which case the sum is XORed into them (incremental encode).
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
    uint32_t accumulate = 0;          // XOR into the destinations instead of overwriting
    const uint32_t *term_index = nullptr; // table term of source t (device, [T]), or null: t
    uint64_t L = 0;                   // symbol length in bytes
    uint64_t count = 0;               // codewords
    // Decode only (either field): the fused coefficients ([codeword][rmax
    // rows][T terms], u16) from which the kernel builds its product tables
    // in LDS; `tables` is then unused.
    const uint16_t *fused = nullptr;
    // Set by the caller: every destination row address and L are multiples
    // of 4, so a launch too small to fill the device may split each
    // codeword's terms into segments whose partial rows are XORed
    // atomically into destinations cleared first (not with `accumulate`).
    uint32_t dst_aligned4 = 0;
    // Split the terms even when the process-wide switch is off (a caller
    // whose launches are always far too small to fill the device).
    uint32_t force_split = 0;
    // GF(2^16) split launches: when set and large enough, every segment
    // stores its partial rows to partials[c][row][seg][L] (row below
    // partial_rows, which bounds every codeword's row count) and a second
    // kernel XORs them into the destinations, instead of atomic XORs into
    // cleared rows.  partials_bytes is the buffer's size.
    uint8_t *partials = nullptr;
    uint32_t partial_rows = 0;
    uint64_t partials_bytes = 0;
    // Set by the launcher.
    uint32_t segs = 1, seg_terms = 0;
};

// Term split of a launch too small to fill the device (see RowsTermsArgs):
// sets segs / seg_terms, true when the launch splits.  Segments are at
// least kSplitMinTerms terms; the split aims at kSplitFillBlocks workgroups.
constexpr uint32_t kSplitMinTerms = 8;
constexpr uint64_t kSplitFillBlocks = 512;
// Target of a forced split (force_split): a lone codeword of tens of
// thousands of terms and a few rows.
#ifndef SCRS_FORCE_SPLIT_BLOCKS
#define SCRS_FORCE_SPLIT_BLOCKS 2048
#endif
constexpr uint64_t kForceSplitFillBlocks = SCRS_FORCE_SPLIT_BLOCKS;
bool plan_term_split(RowsTermsArgs &a, uint64_t active_blocks);
// The process-wide switch of the split (scrs_set_term_split), on by default.
bool term_split_enabled();
// Term segments per codeword that either field's multiply-XOR launcher plans
// for `a` (1 when it does not split): sizes a split launch's partial rows,
// count * partial_rows * segments * L bytes.  Both fields tile 4096 columns
// and 16 rows per workgroup.
uint32_t rows_terms_segments(const RowsTermsArgs &a);
// Whether `a` (split planned) can store partial rows to its buffer.
bool partials_usable(const RowsTermsArgs &a);
// Second kernel of a split launch with partial rows: each destination row
// is the XOR of its segments' partial rows.
hipError_t launch_reduce_partials(const RowsTermsArgs &a, hipStream_t s);
// Clears the destination rows of a split launch (a kernel, no memset).
hipError_t launch_zero_dst_rows(const RowsTermsArgs &a, hipStream_t s);

// Table path: split-nibble product tables staged in LDS, 16-byte loads.
// `vec16` requires every row address and L to be multiples of 16.
hipError_t launch_rows_terms_lds(const RowsTermsArgs &a, bool vec16, hipStream_t s);

// Closed-form Cauchy solve for a batch of codewords (either field).  For
// codeword c with received indices[c][0..k): rows[c] = r (missing data rows),
// missing[c][0..r) ascending, fused[c][j][t] = coefficient of received symbol
// t for missing row j.  coords[i] is the field coordinate of codeword index
// i (x for data, y for repairs).  `work` holds cauchy_solve_work_elems(k,
// m, rmax) u16 per codeword.
//
// With point_wlog set (GF(2^16) only), the generator entry of data index z
// in repair index i is w(z) / (w(i) (coords[i] + coords[z])) for per-index
// weights w, given as discrete logs point_wlog[i] = log w(i).  The solve is
// the unweighted one with fused[j][t] scaled by w(indices[t]) / w(missing j):
// the additive-FFT code's Lagrange generator over the systematic points has
// this form.
struct CauchySolveArgs {
    const uint16_t *indices = nullptr;
    const uint16_t *coords = nullptr;
    uint16_t *fused = nullptr;     // [count][rmax][k]
    uint16_t *missing = nullptr;   // [count][rmax]
    uint32_t *rows = nullptr;      // [count]
    uint16_t *work = nullptr;      // [count][work_stride]
    uint64_t work_stride = 0;
    uint32_t *error_flag = nullptr;
    uint32_t k = 0, m = 0, rmax = 0;
    uint64_t count = 0;
    const uint16_t *gf16_log_exp = nullptr;  // filled by the launcher (GF(2^16))
    const uint16_t *point_wlog = nullptr;    // [k + m] per-index weight logs, or null
};
uint64_t cauchy_solve_work_elems(uint32_t k, uint32_t m, uint32_t rmax);
hipError_t launch_cauchy_solve(const CauchySolveArgs &a, bool gf16, hipStream_t s);

// Product tables and pointer tables from a solve.  GF(2^8) tables use the
// RowsTermsArgs layout (512 B per term and 16-row group); GF(2^16) tables
// use 2048 B per term and group: [q nibble][n][16 rows x 2 bytes].
struct DecodePrepArgs {
    const uint16_t *indices = nullptr;
    const uint16_t *fused = nullptr;
    const uint16_t *missing = nullptr;
    const uint32_t *rows = nullptr;
    const uint8_t *const *payload_ptrs = nullptr;
    const uint8_t *data_base = nullptr;
    const uint8_t *repair_base = nullptr;
    uint8_t *out_base = nullptr;
    uint32_t *tables = nullptr;
    const uint8_t **src_ptrs = nullptr;
    uint8_t **dst_ptrs = nullptr;
    uint8_t **copy_ptrs = nullptr;
    uint32_t k = 0, m = 0, rmax = 0, groups = 0;
    uint64_t L = 0, count = 0;
};
hipError_t launch_decode_prepare(const DecodePrepArgs &a, bool gf16, hipStream_t s);
// Bit-matrix decode tables (kind: kBitmatrixInt8 or kBitmatrixFp4) from a
// GF(2^8) solve (tables only; the pointer tables come from
// launch_decode_prepare with fused = nullptr).
hipError_t launch_bitmatrix_tables(const DecodePrepArgs &a, hipStream_t s, uint32_t kind);

// GF(2^16) multiply-XOR over interleaved little-endian u16 elements, same
// addressing as RowsTermsArgs; tables in the 2048 B per (term, group) layout
// (build_nibble_tables16).
hipError_t launch_rows_terms16(const RowsTermsArgs &a, bool vec16, hipStream_t s);

// Bit-matrix path: the same multiply-XOR as a GF(2) product on the matrix
// cores, with int8 operands (kBitmatrixInt8, 8-term chunks, T <= 64) or
// e2m1 operands (kBitmatrixFp4, 16-term chunks, T <= 256).  Tables hold,
// per codeword, row group g and chunk ch, the MFMA matrix operands of the 8
// output bit planes (8 KiB per chunk): bitmatrix_table_bytes(T, groups, kind)
// bytes per codeword.  Requires L % 4 == 0 and 4-byte aligned rows; no
// accumulate, no term_index.  The kinds equal the C ABI's kernel selectors.
constexpr uint32_t kBitmatrixInt8 = 1, kBitmatrixFp4 = 2;
constexpr uint32_t kBitmatrixMaxChunks = 8;
constexpr uint32_t kFp4MaxChunks = 16;
uint32_t bitmatrix_max_terms(uint32_t kind);
uint64_t bitmatrix_table_bytes(uint32_t T, uint32_t groups, uint32_t kind);
// Host build for a T x R block given term-major (coef[t*R + r]).
void build_bitmatrix_tables(const uint8_t *coef, uint32_t T, uint32_t R, uint32_t groups, uint8_t *out, uint32_t kind);
hipError_t launch_rows_bitmatrix(const RowsTermsArgs &a, hipStream_t s, uint32_t kind);

// Fill `len` bytes with the seeded splitmix64 counter stream: byte i is byte
// (i % 8) of splitmix64 word i / 8 drawn from state seed + (i/8) * golden.
hipError_t launch_fill_splitmix(uint8_t *dst, uint64_t len, uint64_t seed, uint64_t byte_offset,
                                hipStream_t s);

}  // namespace scrs
