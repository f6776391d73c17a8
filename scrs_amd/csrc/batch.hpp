// batch.hpp - one fused launch over a batch of codewords of uneven shape.
//
// A batch is a list of ranges; a range is a run of codewords that share one
// product table (encode jobs of one geometry) or a single decode job with
// its own table.  Work units are (range, codeword chunk, 16-row group,
// column tile); a persistent multiply-XOR kernel pulls units from an atomic
// queue, so GF(2^8) and GF(2^16) jobs of any size mix in one launch.  Decode
// jobs that need a new reconstruction recipe are solved first by one
// persistent launch over the solve descriptors.  All sizes are read from a
// device header, so the same three launches replay from one hipGraph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scrs {

enum BatchRangeFlags : uint32_t {
    kRangeVec16 = 1,   // every row address and L are multiples of 16
    kRangeDecode = 2,  // copies of present data rows (copy pointers)
    kRangeGf16 = 4,    // GF(2^16) tables and elements
    kRangeDst4 = 8,    // destination rows and L are multiples of 4 (term split allowed)
};

struct BatchRange {
    uint64_t L = 0;
    const uint32_t *tables = nullptr;   // T * groups per-(term, group) tables
    const uint8_t *const *src = nullptr;  // cw_count * T source rows
    uint8_t *const *dst = nullptr;        // cw_count * rows destination rows
    uint8_t *const *copy = nullptr;       // cw_count * T copy targets (null entries skip)
    uint32_t T = 0, rows = 0, groups = 0; // groups = ceil(rows / 16), >= 1
    uint32_t tiles = 0;                   // column tiles per codeword
    uint32_t cw_count = 0, ncw = 1;       // codewords, codewords per unit
    uint32_t unit_base = 0;               // first unit of the range
    uint32_t flags = 0;
    // Term segments: with segs > 1 each unit folds seg_terms terms and
    // stores its partial rows to partials[codeword][row][segment], rows of
    // part_stride bytes (L rounded up to 16); a reduce launch XORs them into
    // the destinations.
    uint32_t segs = 1, seg_terms = 0;
    uint8_t *partials = nullptr;
    uint64_t part_stride = 0;
};

// The reduce of one term-split range: destination row (c, row) is the XOR
// of its segments' partial rows, one dword per item.  Items of the launch
// are numbered range after range from item_base.
struct BatchReduce {
    uint8_t *const *dst = nullptr;    // the range's cw_count * rows destination rows
    const uint8_t *partials = nullptr;
    uint64_t L = 0, stride = 0;
    uint64_t item_base = 0;
    uint32_t rows = 0, segs = 0;
};

// A table job: either the closed-form solve of one decode recipe (fused
// coefficients, then tables), or the product tables of an encode geometry
// (coefficient 1 / (x_t + y_row) straight from the coordinates).
struct BatchSolve {
    const uint16_t *coords = nullptr;   // engine coordinate of every wire index
    const uint16_t *idx = nullptr;      // k received indices, received order
    const uint16_t *slot = nullptr;     // k: repair rank of a received repair, 0xFFFF for data
    const uint16_t *missing = nullptr;  // r missing data rows
    uint32_t *work = nullptr;           // 4r + k u32
    uint16_t *fused = nullptr;          // r * k
    uint32_t *tables = nullptr;         // k * groups tables
    uint32_t k = 0, r = 0, groups = 0, gf16 = 0;
    uint32_t encode = 0;                // 1: encode tables for k terms and r = m rows
    uint32_t pad = 0;
};

struct BatchHeader {
    uint32_t nranges = 0, nunits = 0, nsolve = 0, nred = 0;
    // Value of the unit counter when the launch starts (see launch_batch_rows).
    uint32_t claim_base = 0, pad = 0;
    uint64_t red_items = 0;            // dword items of the reduce launch
    const BatchRange *ranges = nullptr;
    const BatchSolve *solves = nullptr;
    const BatchReduce *red = nullptr;  // nred term-split ranges, in item order
};

// Tile widths the multiply-XOR kernel is compiled for (16 bytes per thread).
constexpr uint32_t kBatchTileBytes[3] = {1024, 2048, 4096};

// The launches read every size and array address from the device header, so
// their arguments stay fixed across batches.  The multiply-XOR launch is
// persistent: workgroup b first runs unit b, then draws units from
// counter[0], which only ever grows.  A launch draws exactly `nunits` times
// in all (each workgroup that ran a unit draws once past the end), so the
// caller passes the counter's value at the start as hdr->claim_base and
// adds nunits afterwards: no reset, no memset between launches, and a batch
// with nothing to solve runs the multiply-XOR alone.
hipError_t launch_batch_solve(const BatchHeader *hdr, const uint8_t *gf8_tables, const uint16_t *gf16_log_exp,
                              uint32_t grid, hipStream_t s);
hipError_t launch_batch_rows(const BatchHeader *hdr, uint32_t *counter, uint32_t tile_bytes, uint32_t grid,
                             hipStream_t s);
// XORs the partial rows of the term-split ranges into their destinations.
hipError_t launch_batch_reduce(const BatchHeader *hdr, uint32_t grid, hipStream_t s);
// Persistent grid that fills the device for the given tile width.
uint32_t batch_rows_grid(uint32_t tile_bytes);
// Grid of the reduce launch: enough workgroups for `items` dwords, at most
// kBatchReduceGrid (grid-stride beyond).
constexpr uint32_t kBatchReduceGrid = 2048;
inline uint32_t batch_reduce_grid(uint64_t items) {
    const uint64_t g = (items + 255) / 256;
    return static_cast<uint32_t>(g < 1 ? 1 : (g > kBatchReduceGrid ? kBatchReduceGrid : g));
}

}  // namespace scrs
