// gf256_kernels.hip - GF(2^8) multiply-XOR and decode-solve kernels for gfx950.
//
// The multiply-XOR computes, per byte column p of a codeword,
//   dst_r[p] = sum_t C[t][r] * src_t[p]            (GF(2^8), sum = XOR)
// which is the reference's rows_terms kernel family (scrs/src/simd/rows.rs:257,
// scrs/src/simd/x86.rs:658) applied to a whole batch of codewords at once.
//
// Table path.  For every term t and row group of 16 rows the product table
// holds 32 entries of 16 bytes: entry n (n < 16) is the row vector
// (C[t][r] * n) and entry 16 + n is (C[t][r] * (n << 4)), byte r - 16g of the
// entry for row r.  A source byte x contributes lo[x & 15] ^ hi[x >> 4] to all
// 16 rows at once, so each byte costs two ds_read_b128 and four XORs
// regardless of the row count.  Passes with <= 8 or <= 4 rows use 8- or
// 4-byte entries.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gf_device.hpp"
#include "kernels.hpp"
#include "rows_device.hpp"

namespace scrs {
namespace {

#ifndef SCRS_GF8_AHEAD
#define SCRS_GF8_AHEAD 4     // source terms per load group in the unrolled stream
#endif
constexpr int kThreads = 256;                     // 4 wave64 per workgroup
constexpr int kBytesPerThread = 16;               // one 16-byte load per source
constexpr int kTile = kThreads * kBytesPerThread; // 4096 columns per workgroup
constexpr int kTermChunk = 32;                    // terms staged in LDS at once
constexpr int kRowsPerPass = 16;                  // one 16-byte table entry
constexpr int kGroupDwords = 32 * 4;              // one (term, row group) table
constexpr uint32_t kXcds = 8;                     // MI355X: 8 XCDs, each with its own L2

// LDS for one chunk at 16-byte entries: 32 terms x 32 entries x 16 B = 16 KiB.
constexpr int kLdsBytes = kTermChunk * 32 * 16;

using namespace rowsdev;

// One pass over all terms for up to 16 rows of one codeword.  DEC selects
// the decode addressing: per-codeword source / destination pointer tables
// and verbatim copies of the present data rows; otherwise sources and
// destinations are strided rows of back-to-back codewords (encode).

template <bool VEC, bool DEC, bool FULL, int RW>
__device__ void rows_terms_pass(const RowsTermsArgs &a, uint8_t *lds, uint64_t c, int rg, int rows_here,
                                uint64_t p, bool do_copy, uint32_t tb, uint32_t te) {
    constexpr int D = RW / 4;
    uint32_t acc[16][D];
#pragma unroll
    for (int b = 0; b < 16; b++)
#pragma unroll
        for (int d = 0; d < D; d++) acc[b][d] = 0;

    const bool active = p < a.L;
    const uint32_t *tab_cw = a.tables + c * a.table_cw_stride;
    const uint8_t *src_cw = DEC ? nullptr : a.src_base + c * a.src_cw_stride;
    for (uint32_t t0 = tb; t0 < te; t0 += kTermChunk) {
        const uint32_t tc = min(static_cast<uint32_t>(kTermChunk), te - t0);
        __syncthreads();
        if (DEC && a.fused) {
            // Build this chunk's tables from the fused coefficients instead
            // of reading them: a thread takes one term, one half (entries n
            // or n << 4) and one dword (4 rows), forms the products of the 4
            // packed coefficients with 2^b by byte-wise doubling, and writes
            // the 16 entries as XORs of those (the layout of the staged
            // tables below).
            const uint16_t *fc = a.fused + c * a.rmax * a.T;  // [row][term]
            const uint32_t rows_c = DEC ? load_uniform(a.rows + c) : a.rmax;
            for (uint32_t i = threadIdx.x; i < tc * 2 * D; i += kThreads) {
                const uint32_t t = i / (2 * D), half = (i / D) % 2, d = i % D;
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const uint32_t row = static_cast<uint32_t>(rg) * kRowsPerPass + 4 * d + b;
                    if (row < rows_c) v |= static_cast<uint32_t>(fc[static_cast<uint64_t>(row) * a.T + t0 + t]) << (8 * b);
                }
                if (half) v = gf8_xtime4(gf8_xtime4(gf8_xtime4(gf8_xtime4(v))));
                const uint32_t b[4] = {v, gf8_xtime4(v), gf8_xtime4(gf8_xtime4(v)), gf8_xtime4(gf8_xtime4(gf8_xtime4(v)))};
                uint32_t *dst = reinterpret_cast<uint32_t *>(lds) + t * 32 * D + half * 16 * D + d;
                // Gray-code order: each next entry differs by one basis term.
                uint32_t e = 0;
                dst[0] = 0;
#pragma unroll
                for (int g = 1; g < 16; g++) {
                    e ^= b[__builtin_ctz(g)];
                    dst[(g ^ (g >> 1)) * D] = e;
                }
            }
        } else {
            // Stage this chunk's tables: entry n of term t keeps its first RW bytes.
            for (uint32_t i = threadIdx.x; i < tc * 32 * D; i += kThreads) {
                const uint32_t t = i / (32 * D), rem = i % (32 * D);
                const uint32_t n = rem / D, d = rem % D;
                const uint32_t tt = a.term_index ? a.term_index[t0 + t] : t0 + t;
                reinterpret_cast<uint32_t *>(lds)[i] = tab_cw[(tt * a.groups + rg) * kGroupDwords + n * 4 + d];
            }
        }
        __syncthreads();
        if (!active) continue;
        if (FULL || (VEC && tc == kTermChunk)) {
            // Fully unrolled so every table address is a compile-time LDS
            // offset.  Sources stream in groups of kAhead terms: the next
            // group's loads are in flight while the current group folds.
            constexpr int kAhead = SCRS_GF8_AHEAD;
            constexpr int kGroups = kTermChunk / kAhead;
            const uint32_t p32 = static_cast<uint32_t>(p);
            // Uniform row base + 32-bit lane offset.
            auto row_of = [&](int t) -> const uint8_t * {
                if (DEC) return load_uniform(a.src_ptrs + c * a.T + t0 + t);
                return src_cw + static_cast<uint64_t>(t0 + t) * a.L;
            };
            // Generic (flat) loads on purpose: a flat load also counts
            // against the LDS wait counter, so a wave's first table lookup
            // waits for its prefetched sources.  That throttling measured
            // 1.5% faster on the HBM-bound decode than global loads, and
            // the same on encode.
            auto ld = [&](int t, uint32_t(&w)[4]) {
                const uint4 v = *reinterpret_cast<const uint4 *>(row_of(t) + p32);
                w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
            };
            // Decode writes each present data row verbatim as it streams by.
            // The store's buffer range is L for those terms and zero for
            // received repairs, whose stores the hardware then drops: no
            // per-term branch in the unrolled stream.
            const uint32_t copy_range = do_copy ? static_cast<uint32_t>(a.L) : 0u;
            auto cp = [&](int t, const uint32_t(&w)[4]) {
                if (!DEC) return;
                uint8_t *dst = load_uniform(a.copy_ptrs + c * a.T + t0 + t);
                const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
                    dst ? dst : const_cast<uint8_t *>(row_of(t)), static_cast<short>(0), dst ? copy_range : 0u,
                    0x00020000);
                const v4u32 v = {w[0], w[1], w[2], w[3]};
                __builtin_amdgcn_raw_buffer_store_b128(v, r, p32, 0, 0);
            };
            uint32_t wa[kAhead][4], wb[kAhead][4];
#pragma unroll
            for (int u = 0; u < kAhead; u++) ld(u, wa[u]);
#pragma unroll
            for (int g = 0; g < kGroups; g += 2) {
#pragma unroll
                for (int u = 0; u < kAhead; u++) ld((g + 1) * kAhead + u, wb[u]);
#pragma unroll
                for (int u = 0; u < kAhead; u++) {
                    cp(g * kAhead + u, wa[u]);
                    fold_term<RW>(acc, lds + (g * kAhead + u) * 32 * RW, wa[u]);
                }
                __builtin_amdgcn_sched_barrier(0);
                if (g + 2 < kGroups) {
#pragma unroll
                    for (int u = 0; u < kAhead; u++) ld((g + 2) * kAhead + u, wa[u]);
                }
#pragma unroll
                for (int u = 0; u < kAhead; u++) {
                    cp((g + 1) * kAhead + u, wb[u]);
                    fold_term<RW>(acc, lds + ((g + 1) * kAhead + u) * 32 * RW, wb[u]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            // One term ahead: the next source loads while this one folds.
            auto src_of = [&](uint64_t gt) -> const uint8_t * {
                return DEC ? load_uniform(a.src_ptrs + c * a.T + gt) : src_cw + gt * a.L;
            };
            uint32_t wn[4];
            load16<VEC>(src_of(t0), p, a.L, wn);
#pragma unroll 1
            for (uint32_t t = 0; t < tc; t++) {
                const uint64_t gt = t0 + t;
                const uint32_t w[4] = {wn[0], wn[1], wn[2], wn[3]};
                if (t + 1 < tc) load16<VEC>(src_of(gt + 1), p, a.L, wn);
                if (DEC && do_copy) {
                    uint8_t *cp = load_uniform(a.copy_ptrs + c * a.T + gt);
                    if (cp) store16<VEC>(cp, p, a.L, w);
                }
                fold_term<RW>(acc, lds + t * 32 * RW, w);
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int r = 0; r < RW; r++) {
        if (r >= rows_here) continue;
        const int row = rg * kRowsPerPass + r;
        uint8_t *dst = DEC ? load_uniform(a.dst_ptrs + c * a.rmax + row)
                           : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(row) * a.L;
        const int d = r >> 2, s = (r & 3) * 8;
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            o[j] = ((acc[4 * j + 0][d] >> s) & 0xffu) | (((acc[4 * j + 1][d] >> s) & 0xffu) << 8) |
                   (((acc[4 * j + 2][d] >> s) & 0xffu) << 16) | (((acc[4 * j + 3][d] >> s) & 0xffu) << 24);
        }
        if (a.segs > 1 && a.partials) {  // a term segment: its own partial row
            if (static_cast<uint32_t>(row) >= a.partial_rows) continue;  // outside the buffer's rows
            const uint64_t slot = (c * a.partial_rows + row) * a.segs + tb / a.seg_terms;
            store16<VEC>(a.partials + slot * a.L, p, a.L, o);
            continue;
        }
        if (a.segs > 1) {  // a term segment: XOR into the cleared row
            xor16_atomic(dst, p, a.L, o);
            continue;
        }
        if (!DEC && a.accumulate) {
            uint32_t prev[4];
            load16<VEC>(dst, p, a.L, prev);
#pragma unroll
            for (int j = 0; j < 4; j++) o[j] ^= prev[j];
        }
        store16<VEC>(dst, p, a.L, o);
    }
}

template <bool VEC>
__device__ void copy_only_pass(const RowsTermsArgs &a, uint64_t c, uint64_t p, uint32_t tb, uint32_t te) {
    if (p >= a.L) return;
    for (uint32_t t = tb; t < te; t++) {
        uint8_t *cp = load_uniform(a.copy_ptrs + c * a.T + t);
        if (!cp) continue;
        uint32_t w[4];
        load16<VEC>(load_uniform(a.src_ptrs + c * a.T + t), p, a.L, w);
        store16<VEC>(cp, p, a.L, w);
    }
}

// FULL: the term count is a multiple of the LDS chunk and rows are 16-byte
// aligned, so only the fully unrolled stream is compiled.
template <bool VEC, bool DEC, bool FULL>
__global__ __launch_bounds__(kThreads) void rows_terms_lds_kernel(RowsTermsArgs a, uint32_t tiles, uint32_t units) {
    __shared__ __align__(16) uint8_t lds[kLdsBytes];
    // Workgroups are dealt round-robin to the 8 XCDs.  Give every column
    // tile of one (codeword, row group) unit to the same XCD so the unit's
    // product table is fetched into one L2 only: the slot-th workgroup on
    // XCD x takes unit x + 8 * (slot / tiles), tile slot % tiles.
    const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t unit = xcd + kXcds * (slot / tiles);
    if (unit >= units) return;
    const uint32_t tile = slot % tiles;
    // Unit = (codeword, row group, term segment).
    const uint32_t seg = unit % a.segs, cg = unit / a.segs;
    const uint64_t c = cg / a.groups_launched;
    const int rg = static_cast<int>(cg % a.groups_launched);
    const uint32_t tb = seg * a.seg_terms, te = min(a.T, tb + a.seg_terms);
    const uint32_t R = DEC ? load_uniform(a.rows + c) : a.rmax;
    const uint32_t base = static_cast<uint32_t>(rg) * kRowsPerPass;
    const int rows_here = static_cast<int>(R > base ? min(static_cast<uint32_t>(kRowsPerPass), R - base) : 0u);
    const bool do_copy = DEC && rg == 0;
    if (rows_here == 0 && !do_copy) return;
    const uint64_t p = static_cast<uint64_t>(tile) * kTile + threadIdx.x * kBytesPerThread;
    if (DEC && rows_here == 0) {
        copy_only_pass<VEC>(a, c, p, tb, te);
    } else if (rows_here <= 4) {
        rows_terms_pass<VEC, DEC, FULL, 4>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
    } else if (rows_here <= 8) {
        rows_terms_pass<VEC, DEC, FULL, 8>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
    } else {
        rows_terms_pass<VEC, DEC, FULL, 16>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
    }
}

// Destination rows of a term-split launch start from zero: one workgroup
// per (row, codeword); rows are dword-aligned (the split requires it).
__global__ __launch_bounds__(256) void zero_dst_rows_kernel(RowsTermsArgs a) {
    const uint64_t c = blockIdx.y;
    const uint32_t r = blockIdx.x;
    const bool dec = a.dst_ptrs != nullptr;
    if (dec && r >= load_uniform(a.rows + c)) return;
    uint8_t *row = dec ? load_uniform(a.dst_ptrs + c * a.rmax + r)
                       : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(r) * a.L;
    for (uint64_t i = threadIdx.x; i < a.L / 4; i += blockDim.x) *gbl_mut<uint32_t>(row + 4 * i) = 0;
}

__global__ void fill_splitmix_kernel(uint8_t *dst, uint64_t len, uint64_t seed, uint64_t word0) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t w = g; w * 8 < len; w += stride) {
        uint64_t z = seed + (word0 + w + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        if (w * 8 + 8 <= len) {
            *reinterpret_cast<uint64_t *>(dst + w * 8) = z;
        } else {
            for (uint64_t b = 0; w * 8 + b < len; b++) dst[w * 8 + b] = static_cast<uint8_t>(z >> (8 * b));
        }
    }
}

}  // namespace

bool plan_term_split(RowsTermsArgs &a, uint64_t active_blocks) {
    a.segs = 1;
    a.seg_terms = a.T;
    const uint64_t fill = a.force_split ? kForceSplitFillBlocks : kSplitFillBlocks;
    if (!(term_split_enabled() || a.force_split) || a.accumulate || !a.dst_aligned4 || a.L % 4 || a.rmax == 0 ||
        a.T < 2 * kSplitMinTerms ||
        active_blocks >= fill || a.count > 65535)
        return false;
    const uint64_t want = std::min<uint64_t>(a.T / kSplitMinTerms, (fill + active_blocks - 1) / active_blocks);
    if (want < 2) return false;
    a.seg_terms = static_cast<uint32_t>((a.T + want - 1) / want);
    // A forced split keeps whole chunks of 16 (or 8) terms per segment when
    // the term count allows, so the segments run the unrolled kernels.
    const uint32_t unit = a.T % 16 == 0 ? 16u : (a.T % 8 == 0 ? 8u : 1u);
    if (a.force_split) a.seg_terms = (a.seg_terms + unit - 1) / unit * unit;
    a.segs = (a.T + a.seg_terms - 1) / a.seg_terms;
    if (a.segs < 2) {
        a.segs = 1;
        a.seg_terms = a.T;
        return false;
    }
    return true;
}

hipError_t launch_zero_dst_rows(const RowsTermsArgs &a, hipStream_t s) {
    if (a.count == 0 || a.rmax == 0 || a.count > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_dst_rows_kernel, dim3(a.rmax, static_cast<uint32_t>(a.count)), dim3(256), 0, s, a);
    return hipGetLastError();
}

uint32_t rows_terms_segments(const RowsTermsArgs &args) {
    RowsTermsArgs a = args;
    if (a.count == 0 || a.L == 0) return 1;
    const uint64_t tiles = (a.L + kTile - 1) / kTile;
    const uint32_t groups = a.rmax ? (a.rmax + kRowsPerPass - 1) / kRowsPerPass : 1;
    plan_term_split(a, a.count * groups * tiles);
    return a.segs;
}

bool partials_usable(const RowsTermsArgs &a) {
    return a.segs > 1 && a.partials && a.partial_rows != 0 && a.partial_rows <= a.rmax &&
           a.count * a.partial_rows * a.segs * a.L <= a.partials_bytes && (a.L / 4 + 255) / 256 <= 0x7fffffffull &&
           a.count <= 65535;
}

hipError_t launch_rows_terms_lds(const RowsTermsArgs &args, bool vec16, hipStream_t s) {
    // The 16-byte paths address a row with a 32-bit column offset.
    if (args.L > 0xffffffffull) vec16 = false;
    if (args.count == 0 || args.L == 0) return hipSuccess;
    RowsTermsArgs a = args;
    const uint64_t tiles = (a.L + kTile - 1) / kTile;
    const uint32_t groups = a.rmax ? (a.rmax + kRowsPerPass - 1) / kRowsPerPass : 1;
    if (groups > a.groups + (a.rmax == 0)) return hipErrorInvalidValue;
    a.groups_launched = groups;
    // A launch too small to fill the device splits every codeword's terms.
    const bool split = plan_term_split(a, a.count * groups * tiles);
    const uint64_t units = a.count * groups * a.segs;
    const uint64_t blocks = (units + kXcds - 1) / kXcds * kXcds * tiles;
    if (blocks > 0x7fffffffull || units > 0xffffffffull) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(blocks));
    const uint32_t t32 = static_cast<uint32_t>(tiles), u32 = static_cast<uint32_t>(units);
    const bool dec = a.src_ptrs != nullptr;
    if (dec != (a.dst_ptrs != nullptr) || dec != (a.copy_ptrs != nullptr) || dec != (a.rows != nullptr))
        return hipErrorInvalidValue;
    const bool full = vec16 && a.T % kTermChunk == 0 && !split;
    // Partial rows and a reduce when the caller's buffer holds them, else
    // atomic XORs into rows cleared first.
    if (!partials_usable(a)) a.partials = nullptr;
    if (split && !a.partials) {
        hipError_t err = launch_zero_dst_rows(a, s);
        if (err != hipSuccess) return err;
    }
    if (full && dec)
        hipLaunchKernelGGL((rows_terms_lds_kernel<true, true, true>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (full)
        hipLaunchKernelGGL((rows_terms_lds_kernel<true, false, true>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (vec16 && dec)
        hipLaunchKernelGGL((rows_terms_lds_kernel<true, true, false>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (vec16)
        hipLaunchKernelGGL((rows_terms_lds_kernel<true, false, false>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (dec)
        hipLaunchKernelGGL((rows_terms_lds_kernel<false, true, false>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else
        hipLaunchKernelGGL((rows_terms_lds_kernel<false, false, false>), grid, dim3(kThreads), 0, s, a, t32, u32);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || !a.partials) return err;
    return launch_reduce_partials(a, s);
}

hipError_t launch_fill_splitmix(uint8_t *dst, uint64_t len, uint64_t seed, uint64_t byte_offset, hipStream_t s) {
    if (len == 0) return hipSuccess;
    if (byte_offset % 8) return hipErrorInvalidValue;
    const uint64_t words = (len + 7) / 8;
    const uint64_t want = (words + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(fill_splitmix_kernel, dim3(blocks), dim3(256), 0, s, dst, len, seed, byte_offset / 8);
    return hipGetLastError();
}

}  // namespace scrs
