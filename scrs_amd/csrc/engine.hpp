// engine.hpp - device decode engine and coefficient plans shared by the
// batch decoder, the streaming decoder and the incremental encoder.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/scrs.h"
#include "device.hpp"
#include "kernels.hpp"

namespace scrs {

inline bool engine_is_gf256(int32_t engine) {
    return engine == SCRS_ENGINE_STANDARD_CAUCHY || engine == SCRS_ENGINE_GOOD_CAUCHY;
}

size_t engine_capacity(int32_t engine);

// Field coordinates of every codeword index: data i -> x_i, repair j -> y_j.
// Standard Cauchy x_i = i, y_j = k + j (matrices/cauchy.rs:88-99); Good
// Cauchy 3^i, 3^(k+j) (matrices/good_cauchy.rs:59-64); Tower g^i, g^(k+j)
// with g = 0x0108 (tower/cauchy.rs:46-63).
std::vector<uint16_t> cauchy_coordinates(const scrs_profile &p);

// Product-table bytes per (term, 16-row group).
inline uint64_t table_group_bytes(bool gf16) { return gf16 ? 2048 : 512; }

// k x m coefficient block of a Cauchy engine as a table for `T` terms, all
// m rows, one group layout per field.  `terms` lists the source rows (data
// indices) in term order.
std::vector<uint32_t> coefficient_tables(const scrs_profile &p, const std::vector<uint32_t> &terms,
                                         uint32_t *groups_out);

// Launch the multiply-XOR of the profile's field.
hipError_t launch_rows_terms_field(bool gf16, const RowsTermsArgs &a, bool vec16, hipStream_t s);

// Partial-row buffer of a codec's term-split launches.  A launch too small
// to fill the device splits each codeword's terms into segments; each
// segment stores its partial rows here and a reduce kernel XORs them into
// the destinations.  One launch holds the buffer at a time (the lock), and a
// launch on another stream first waits for the previous one's work.
struct SplitBuffer {
    std::mutex mu;
    DevBuf buf;
    hipEvent_t done = nullptr;
    static constexpr uint64_t kMaxBytes = 256ull << 20;

    SplitBuffer() = default;
    SplitBuffer(const SplitBuffer &) = delete;
    SplitBuffer &operator=(const SplitBuffer &) = delete;
    ~SplitBuffer();
    // When `a`'s launch splits, give it partial rows for rows < `rows` (an
    // upper bound of every codeword's row count): takes the lock, sizes the
    // buffer (only if may_allocate) and orders `s` after the buffer's previous
    // launch.  True when attached; the caller launches and then calls
    // release(s).  False leaves `a` to the launcher's atomic split.
    bool acquire(RowsTermsArgs &a, uint32_t rows, hipStream_t s, bool may_allocate);
    void release(hipStream_t s);
};

// Device workspace of one decode launch: received indices, the solve's
// outputs and work, and the pointer tables, for `cap` codewords.
struct DecodeWorkspace {
    uint16_t *idx = nullptr;      // [cap][k]
    uint16_t *fused = nullptr;    // [cap][rmax][k]
    uint16_t *missing = nullptr;  // [cap][rmax]
    uint32_t *rows = nullptr;     // [cap]
    uint16_t *work = nullptr;     // [cap][cauchy_solve_work_elems]
    const uint8_t **src = nullptr;  // [cap][k]
    uint8_t **dst = nullptr;        // [cap][rmax]
    uint8_t **copy = nullptr;       // [cap][k]
    uint32_t *tables = nullptr;     // [cap][cw_table_bytes / 4], or null when no kernel reads them
    uint64_t cap = 0;
};

// Solve + tables + pointers + multiply-XOR for batches of codewords of one
// Cauchy geometry (either field).
struct DecodeEngine {
    scrs_profile prof{};
    bool gf16 = false;
    uint32_t k = 0, m = 0, rmax = 0, groups = 0;
    uint64_t L = 0;
    DevBuf d_coords;
    DevBuf d_idx, d_fused, d_missing, d_rows, d_work, d_flag, d_tables, d_src, d_dst, d_copy;
    HostBuf h_idx;
    hipEvent_t idx_copied = nullptr;
    uint64_t cap = 0;     // codewords the workspace holds
    uint64_t chunk = 0;   // codewords per launch (bounded table memory)
    // GF(2^8) table path: write the product tables to d_tables even though
    // the multiply-XOR builds its own in LDS (a recipe cache copies them).
    bool write_tables = false;
    // Weighted solve (CauchySolveArgs::point_wlog) and forced term split:
    // the additive-FFT code's direct decode of a few erasures.
    DevBuf d_wlog;
    bool weighted = false, force_split = false;
    // Forced splits: the caller's exact row count of the next enqueue (0 =
    // unknown), which bounds the partial rows; other launches bound them by
    // rmax.
    uint32_t rows_hint = 0;
    SplitBuffer split;

    DecodeEngine() = default;
    DecodeEngine(const DecodeEngine &) = delete;
    DecodeEngine &operator=(const DecodeEngine &) = delete;
    ~DecodeEngine();

    // Returns a construction status (TooManySymbols{cap} on device OOM).
    scrs_status init(const scrs_profile &p);
    // The additive-FFT code's direct decode of at most `max_missing` data
    // erasures: points are the wire indices, and the generator entry of data
    // z in repair i is P(i) / ((i + z) P'(z)) for P the product of (x + d)
    // over the data points d < k; wlog[x] (x < k + m) = log(1 / P'(x)) for
    // data and log(1 / P(x)) for repairs.
    scrs_status init_afft_direct(const scrs_profile &p, uint32_t max_missing, const std::vector<uint16_t> &wlog);
    // Shared tail of the two: coordinates, events, flag, launch chunking.
    scrs_status init_device(const std::vector<uint16_t> &coords);
    // Enqueue the decode of `count` codewords whose indices (host memory,
    // count x k, received order) are already validated.  Received payloads
    // come from `payload_ptrs` (device array, count x k) or from the
    // back-to-back data / repair layout.  Output rows: out + c*k*L + i*L.
    // `bitmatrix` (kBitmatrixInt8 or kBitmatrixFp4, 0 for the tables) runs
    // the GF(2^8) multiply-XOR on that bit-matrix kernel (payload rows and
    // outputs 4-byte aligned, k <= bitmatrix_max_terms).
    scrs_status enqueue(const uint16_t *indices, const uint8_t *const *payload_ptrs, const uint8_t *data_base,
                        const uint8_t *repair_base, uint8_t *out, uint64_t count, hipStream_t s, bool vec16,
                        uint32_t bitmatrix = 0);
    // The chain of one launch chunk of `cn` codewords whose indices are
    // already in ws.idx (payload, data, repair and output bases of the
    // chunk's first codeword).
    // may_allocate = false: a split launch uses the partial-row buffer only
    // if it already holds the launch, else atomics.
    scrs_status run_chunk(const DecodeWorkspace &ws, const uint8_t *const *payload_ptrs, const uint8_t *data_base,
                          const uint8_t *repair_base, uint8_t *out, uint64_t cn, hipStream_t s, bool vec16,
                          uint32_t bitmatrix, bool may_allocate = true);
    // Bytes a caller-provided workspace needs for `count` codewords of the
    // GF(2^8) table kernel (no product tables in HBM): each array starts on
    // a 256-byte boundary.  carve_workspace lays them out from `base`.
    uint64_t workspace_bytes(uint64_t count) const;
    DecodeWorkspace carve_workspace(uint8_t *base, uint64_t count) const;
    // Decode one codeword with a reconstruction recipe computed earlier (its
    // product tables, missing rows and row count on the device): only the
    // pointer tables and the multiply-XOR run.
    scrs_status enqueue_recipe(const uint16_t *indices, const uint32_t *tables, const uint16_t *missing,
                               const uint32_t *rows, const uint8_t *data_base, const uint8_t *repair_base,
                               uint8_t *out, hipStream_t s, bool vec16);
    // Table bytes per codeword of the table kernel (0) or a bit-matrix kernel.
    uint64_t cw_table_bytes(uint32_t bitmatrix) const {
        return bitmatrix ? bitmatrix_table_bytes(k, groups, bitmatrix)
                         : static_cast<uint64_t>(k) * groups * table_group_bytes(gf16);
    }
    // Bytes of one codeword's recipe: tables, missing rows, row count.
    uint64_t recipe_table_bytes() const { return static_cast<uint64_t>(k) * groups * table_group_bytes(gf16); }
    // Non-zero if a solve found an inconsistent submatrix since the last reset.
    hipError_t read_error(uint32_t *flag, hipStream_t s);
    hipError_t clear_error(hipStream_t s);
};

}  // namespace scrs
