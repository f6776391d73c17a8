// afft.hpp - device additive FFT over GF(2^16): kernels and the codec engine.
//
// Symbols are rows of L bytes holding L/2 little-endian GF(2^16) elements;
// every element column is an independent transform over the rows.  A pass
// loads a group of 2^g rows (the rows that differ only in g consecutive bits
// of the row index) for a 256-byte column tile into registers and LDS, runs
// its butterfly levels and writes it back.  A transform over 2^D rows takes
// ceil(D / 8) passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/scrs.h"
#include "device.hpp"
#include "engine.hpp"

namespace scrs {

enum AfftPassFlags : uint32_t {
    kAfftInv = 1,    // inverse levels over the pass bits (bottom-up)
    kAfftDeriv = 2,  // formal derivative contributions of the pass bits
    kAfftFwd = 4,    // forward levels over the pass bits (top-down)
    // The derivative with every factor 1: plain XORs of the rows.  With the
    // coefficients scaled by T(q) = prod over bits b of q of d_b before and
    // the sums by 1/T(p) after, it equals the derivative with the factors d_b:
    // sum over b clear in p of d_b C[p | 2^b] = (1/T(p)) sum of T(p | 2^b) C[p | 2^b].
    kAfftDerivUnit = 8,
};

struct AfftPassArgs {
    // Input rows: in_rows[r] (null = zero row), else in_base + r*L for
    // r < in_limit and zero above.  Codeword z adds z*in_cw_stride.
    const uint8_t *const *in_rows = nullptr;
    const uint8_t *in_base = nullptr;
    uint64_t in_limit = 0, in_cw_stride = 0;
    // Derivative accumulator XORed into the derivative (rows of aux_base).
    const uint8_t *aux_base = nullptr;
    uint64_t aux_cw_stride = 0;
    // Output rows: out_rows[r] (null = not stored), else out_base +
    // (r - out_first)*L for out_first <= r < out_limit.
    uint8_t *const *out_rows = nullptr;
    uint8_t *out_base = nullptr;
    uint64_t out_first = 0, out_limit = 0, out_cw_stride = 0;
    // Loaded rows are also copied, unscaled, to copy_rows[r] when non-null.
    uint8_t *const *copy_rows = nullptr;
    // Per-row constant factors (row r's at [r]) multiplied in after the
    // load / before the store; the pass builds their product tables in LDS.
    const uint16_t *scale_in = nullptr;
    const uint16_t *scale_out = nullptr;
    // Per-row factors multiplied into the derivative (after the
    // accumulator, before the forward levels).
    const uint16_t *scale_der = nullptr;
    const uint16_t *node_fac = nullptr;    // 65536 node factors (the pass's node tables are built from them)
    const uint16_t *deriv_tabs = nullptr;  // 16 derivative tables
    uint64_t L = 0;
    uint32_t b_lo = 0, g = 0;              // pass bits [b_lo, b_lo + g)
    uint32_t inv_base = 0, fwd_base = 0;   // point of buffer row 0, per direction
    uint32_t flags = 0;
    uint32_t row_limit = 0;                // only rows < row_limit are needed downstream
    // Every pointer in in_rows / copy_rows / out_rows is a multiple of 4.
    bool rows_aligned4 = false;
};

// Launch one pass over all 2^D rows (groups = 2^(D-g)) for `count`
// codewords.
hipError_t launch_afft_pass(const AfftPassArgs &a, uint32_t D, uint32_t count, hipStream_t s);
size_t afft_pass_lds_bytes(uint32_t g, uint32_t flags, bool scale_in, bool scale_out, bool scale_der,
                           bool wide = false, bool mixed = false);

// Nibble product tables: out[i*64 + q*16 + n] = c[i] * (n << 4q).
hipError_t launch_afft_tables(const uint16_t *c, uint32_t count, uint16_t *out, hipStream_t s);
// Wide tables: out[i*128 + n] = c[i]*n for n < 64, out[i*128 + 64 + n] =
// c[i]*(n << 6) and out[i*128 + 96 + n] = c[i]*(n << 11) for n < 32.

// Locator constants for the point set given by the bitmap `known_bits`
// (bit x of word x/32) over [0, 2^D): the per-row scale of the decode.
struct AfftLocatorArgs {
    const uint32_t *known_bits = nullptr;
    const uint32_t *log_wht = nullptr;  // WHT of log[x], 2^D entries
    const uint16_t *exp = nullptr;      // 65535 entries
    uint32_t *work = nullptr;           // 2^D u32
    uint16_t *out = nullptr;            // 2^D constants
    uint32_t D = 0;
};
hipError_t launch_afft_locator(const AfftLocatorArgs &a, hipStream_t s);

// Row pointer tables of one decode over [0, 2^D): known rows read from
// source (layout: x < k from data_base + x*L, else repair_base + (x-k)*L),
// present data copied to out, missing data written to out.
struct AfftRowsArgs {
    const uint32_t *known_bits = nullptr;
    const uint8_t *data_base = nullptr;
    const uint8_t *repair_base = nullptr;
    uint8_t *out = nullptr;
    const uint8_t **in_rows = nullptr;
    uint8_t **copy_rows = nullptr;
    uint8_t **out_rows = nullptr;
    uint32_t k = 0, n = 0, D = 0;
    uint64_t L = 0;
};
hipError_t launch_afft_rows(const AfftRowsArgs &a, hipStream_t s);
// in_rows[indices[t]] = ptrs[t] for t < count.
hipError_t launch_afft_scatter_ptrs(const uint16_t *indices, const uint8_t *const *ptrs, uint32_t count,
                                    const uint8_t **in_rows, hipStream_t s);

// Per-device constant tables.
struct AfftDeviceTables {
    DevBuf deriv_tabs, exp;
    // The factor of every tree node (65536 u16), from which the pass kernel
    // builds its node tables in LDS.
    DevBuf node_fac;
    // T(q) = prod over the bits b of q of the derivative factor d_b, and
    // 1/T(q), for q < 65536 (kAfftDerivUnit).
    DevBuf deriv_scale, deriv_unscale;
};
hipError_t afft_device_tables(const AfftDeviceTables **out);

// Decodes with at most this many missing data symbols take the direct path
// (the reference's targeted path, afft/decoder.rs:360-441): one weighted
// closed-form solve and one multiply-XOR of the k received rows into the
// missing ones, instead of the transforms over the whole domain.  The
// decoded bytes are the same on either path.  kAfftDirectCap bounds the
// process-wide setting (scrs_afft_set_targeted_max).
constexpr uint32_t kAfftDirectCap = 64;
constexpr uint32_t kAfftTargetedDefault = 5;
uint32_t afft_targeted_max();

// Reed-Solomon codec over GF(2^16) evaluated at the points 0..n-1
// (points = wire indices, afft/profile.rs:41-46): data symbol i is P(i),
// repair j is P(k + j), deg P < k.
struct AfftEngine {
    scrs_profile prof{};
    uint32_t k = 0, m = 0, n = 0, D = 0, Dk = 0;  // 2^D >= n, 2^Dk >= k
    uint64_t N = 0, L = 0;
    bool fused = false;     // k = 2^Dk, m <= k: interpolate then evaluate the coset [k, 2k)
    const AfftDeviceTables *dev = nullptr;
    DevBuf enc_scale;       // systematic locator factors (general encode)
    DevBuf log_wht;         // decode locator input
    DevBuf work, work2;
    DevBuf bits, loc_work, loc_out, in_rows, copy_rows, out_rows;
    std::vector<uint32_t> h_bits;
    HostBuf h_bits_pinned;
    hipEvent_t bits_copied = nullptr;
    uint64_t enc_cap = 0;   // codewords the encode workspace holds
    DecodeEngine direct;    // decode of a few erasures (afft_targeted_max)
    // The workspaces serve one call at a time: calls lock `mu`, and work
    // enqueued on one stream waits for the previous call's work on another.
    std::mutex mu;
    hipEvent_t last = nullptr;

    AfftEngine() = default;
    AfftEngine(const AfftEngine &) = delete;
    AfftEngine &operator=(const AfftEngine &) = delete;
    ~AfftEngine();

    scrs_status init(const scrs_profile &p, bool for_decode);
    // Order `s` after the previous call (under `mu`); `finish` records it.
    hipError_t begin(hipStream_t s) { return hipStreamWaitEvent(s, last, 0); }
    hipError_t finish(hipStream_t s) { return hipEventRecord(last, s); }
    // `count` codewords back to back: data c at data + c*k*L, repairs at
    // repairs + c*m*L.
    scrs_status encode(const uint8_t *data, uint8_t *repairs, uint64_t count, hipStream_t s);
    // One codeword.  `indices` lists the received wire indices (host,
    // exactly k, distinct, validated); received64, when given, is the same
    // set as a bitmap over the n indices (bit i of word i / 64).  Payloads
    // from `ptrs` (device array, aligned with `indices`) or from the data /
    // repair layout.
    scrs_status decode(const uint16_t *indices, const uint8_t *const *ptrs, const uint8_t *data_base,
                       const uint8_t *repair_base, uint8_t *out, hipStream_t s,
                       const uint64_t *received64 = nullptr);
};

}  // namespace scrs
