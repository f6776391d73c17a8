// gf16_kernels.hip - GF(2^16) Tower multiply-XOR for gfx950.
//
// Symbols are interleaved little-endian u16 field elements
// (scrs/src/algebra/gf65536.rs:53-60).  The multiply-XOR
//   dst_r ^= sum_t C[t][r] * src_t
// is the reference's tower AXPY family (scrs/src/tower/payload.rs:7,158,433)
// over a batch of codewords.  Multiplication by a constant is linear over
// GF(2), so c * x = sum over the four nibbles q of x of c * (n_q << 4q).  The
// product table of term t and 16-row group g holds, for every nibble
// position q and value n, the 16 row products (2 bytes each): one element
// costs four 32-byte LDS lookups and eight XOR3s for all 16 rows.
#include <hip/hip_runtime.h>

#include "gf_device.hpp"
#include "kernels.hpp"
#include "rows_device.hpp"

namespace scrs {
namespace {

#ifndef SCRS_GF16_AHEAD
#define SCRS_GF16_AHEAD 2    // source terms per load group in the unrolled stream
#endif
constexpr int kThreads = 256;
constexpr int kTile = kThreads * 16;          // bytes per workgroup tile
constexpr int kTermChunk = 16;                // terms staged per LDS fill
constexpr int kRowsPerPass = 16;
constexpr int kGroupBytes = 4 * 16 * 32;      // one (term, group) table: 2 KiB
constexpr uint32_t kXcds = 8;

using namespace rowsdev;

// FULL (8 or 16): the term count is a multiple of FULL, rows are 16-byte
// aligned and the terms run FULL at a time fully unrolled; 0: a rolled loop
// over chunks of kTermChunk terms.
template <bool VEC, bool DEC, int RW, int FULL>
__device__ void pass16(const RowsTermsArgs &a, uint8_t *lds, uint64_t c, int rg, int rows_here, uint64_t p,
                       bool do_copy, uint32_t tb, uint32_t te) {
    constexpr int EB = 2 * RW;
    constexpr int D = EB / 4;
    uint32_t acc[8][D];
#pragma unroll
    for (int e = 0; e < 8; e++)
#pragma unroll
        for (int d = 0; d < D; d++) acc[e][d] = 0;
    const bool active = p < a.L;
    const uint8_t *tab_cw = reinterpret_cast<const uint8_t *>(a.tables + c * a.table_cw_stride);
    const uint8_t *src_cw = DEC ? nullptr : a.src_base + c * a.src_cw_stride;
    constexpr uint32_t kChunk = FULL ? FULL : kTermChunk;
    for (uint32_t t0 = tb; t0 < te; t0 += kChunk) {
        const uint32_t tc = min(kChunk, te - t0);
        __syncthreads();
        if (DEC && a.fused) {
            __builtin_amdgcn_sched_barrier(0);
            // Build this chunk's tables from the fused coefficients instead
            // of reading them: a thread takes one term, one row pair (dword
            // d) and one nibble position q, forms the products of the two
            // packed coefficients with the four basis elements of q, and
            // writes the 16 entries of q as XORs of those.  For bits b < 8
            // the basis element 2^b is the base-field x^b, which doubles
            // both bytes of a + b u; for b >= 8 it is x^(b-8) u, and
            // (a + b u) u = 0x20 b + (a + b) u (u^2 = u + 0x20).
            const uint16_t *fc = a.fused + c * a.rmax * a.T;  // [row][term]
            const uint32_t rows_c = load_uniform(a.rows + c);
            // The thread's task indices are recomputed per chunk: hoisted out
            // of the term loop they would stay live across the fold.
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            for (uint32_t i = tid; i < tc * D * 4; i += kThreads) {
                const uint32_t q = i & 3u, d = (i >> 2) % D, t = (i >> 2) / D;
                const uint32_t row = static_cast<uint32_t>(rg) * kRowsPerPass + 2 * d;
                const uint64_t col = t0 + t;
                const uint32_t f0 = row < rows_c ? fc[static_cast<uint64_t>(row) * a.T + col] : 0u;
                const uint32_t f1 = row + 1 < rows_c ? fc[static_cast<uint64_t>(row + 1) * a.T + col] : 0u;
                uint32_t v = f0 | (f1 << 16);
                if (q >= 2) {
                    const uint32_t lo = v & 0x00ff00ffu, hi = (v >> 8) & 0x00ff00ffu;
                    uint32_t hi20 = hi;
#pragma unroll
                    for (int x = 0; x < 5; x++) hi20 = gf8_xtime4(hi20);
                    v = hi20 | ((lo ^ hi) << 8);
                }
                if (q & 1u) v = gf8_xtime4(gf8_xtime4(gf8_xtime4(gf8_xtime4(v))));
                uint32_t b[4];
                b[0] = v;
#pragma unroll
                for (int x = 1; x < 4; x++) b[x] = gf8_xtime4(b[x - 1]);
                uint32_t *dst = reinterpret_cast<uint32_t *>(lds) + t * 64 * D;
                // Gray-code order: each next entry differs by one basis term.
                uint32_t e = 0;
                dst[lds16_index<EB>(q * 16, d)] = 0;
#pragma unroll
                for (int g = 1; g < 16; g++) {
                    e ^= b[__builtin_ctz(g)];
                    dst[lds16_index<EB>(q * 16 + (g ^ (g >> 1)), d)] = e;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        } else {
            // Stage: entry (q, n) of term t keeps its first EB bytes (rows 0..RW-1).
            for (uint32_t i = threadIdx.x; i < tc * 64 * D; i += kThreads) {
                const uint32_t t = i / (64 * D), rem = i % (64 * D);
                const uint32_t e = rem / D, d = rem % D;
                const uint32_t tt = a.term_index ? a.term_index[t0 + t] : t0 + t;
                reinterpret_cast<uint32_t *>(lds)[t * 64 * D + lds16_index<EB>(e, d)] =
                    *reinterpret_cast<const uint32_t *>(tab_cw + (static_cast<uint64_t>(tt) * a.groups + rg) * kGroupBytes +
                                                        e * 32 + d * 4);
            }
        }
        __syncthreads();
        if (!active) continue;
        if constexpr (FULL != 0) {
            // The chunk is unrolled so that every table address is a
            // compile-time LDS offset, and sources stream in groups of
            // kAhead terms with the next group's loads in flight.
            constexpr int kAhead = SCRS_GF16_AHEAD;
            constexpr int kGroups = FULL / kAhead;
            static_assert(FULL % (2 * kAhead) == 0, "groups run in pairs");
            const uint32_t p32 = static_cast<uint32_t>(p);
            auto row_of = [&](int t) -> const uint8_t * {
                if (DEC) return load_uniform(a.src_ptrs + c * a.T + t0 + t);
                return src_cw + static_cast<uint64_t>(t0 + t) * a.L;
            };
            auto ld = [&](int t, uint32_t(&w)[4]) {
                const v4u32 v = *gbl<v4u32>(row_of(t) + p32);
                w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
            };
            // Present data rows are written verbatim as they stream by: the
            // store's buffer range is L for those terms and zero otherwise,
            // so the hardware drops the other stores without a branch.
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
                    fold_term16<EB>(acc, lds + (g * kAhead + u) * 64 * EB, wa[u]);
                }
                __builtin_amdgcn_sched_barrier(0);
                if (g + 2 < kGroups) {
#pragma unroll
                    for (int u = 0; u < kAhead; u++) ld((g + 2) * kAhead + u, wa[u]);
                }
#pragma unroll
                for (int u = 0; u < kAhead; u++) {
                    cp((g + 1) * kAhead + u, wb[u]);
                    fold_term16<EB>(acc, lds + ((g + 1) * kAhead + u) * 64 * EB, wb[u]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            continue;
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
            fold_term16<EB>(acc, lds + t * 64 * EB, w);
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
        const int d = r >> 1, sh = (r & 1) * 16;
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            o[j] = ((acc[2 * j][d] >> sh) & 0xffffu) | (((acc[2 * j + 1][d] >> sh) & 0xffffu) << 16);
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
        if (a.accumulate) {
            uint32_t prev[4];
            load16<VEC>(dst, p, a.L, prev);
#pragma unroll
            for (int j = 0; j < 4; j++) o[j] ^= prev[j];
        }
        store16<VEC>(dst, p, a.L, o);
    }
}

template <bool VEC>
__device__ void copy_only16(const RowsTermsArgs &a, uint64_t c, uint64_t p, uint32_t tb, uint32_t te) {
    if (p >= a.L) return;
    for (uint32_t t = tb; t < te; t++) {
        uint8_t *cp = load_uniform(a.copy_ptrs + c * a.T + t);
        if (!cp) continue;
        uint32_t w[4];
        load16<VEC>(load_uniform(a.src_ptrs + c * a.T + t), p, a.L, w);
        store16<VEC>(cp, p, a.L, w);
    }
}

// Second kernel of a split launch with partial rows: dword d of row `row`
// of codeword blockIdx.z is the XOR of that dword over the segments.
__global__ __launch_bounds__(256) void reduce_partials_kernel(RowsTermsArgs a) {
    const uint64_t c = blockIdx.z;
    const uint32_t row = blockIdx.y;
    const uint64_t d = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool dec = a.dst_ptrs != nullptr;
    if (d >= a.L / 4 || (dec && row >= load_uniform(a.rows + c))) return;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.partials) +
                          (c * a.partial_rows + row) * a.segs * (a.L / 4) + d;
    uint32_t v = 0;
#pragma unroll 8
    for (uint32_t sg = 0; sg < a.segs; sg++) v ^= *gbl<uint32_t>(src + static_cast<uint64_t>(sg) * (a.L / 4));
    uint8_t *dst = dec ? load_uniform(a.dst_ptrs + c * a.rmax + row)
                       : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(row) * a.L;
    *gbl_mut<uint32_t>(dst + 4 * d) = v;
}

// The same for many segments over few dwords (too few threads above to
// fill the device, each folding thousands of partials in sequence): a
// workgroup covers 4 dwords with 64 lanes per dword, each lane XORing every
// 64th segment's partial, and the lanes combine in LDS.
constexpr uint32_t kReduceDwords = 4, kReduceLanes = 64;
__global__ __launch_bounds__(256) void reduce_partials_wide_kernel(RowsTermsArgs a) {
    __shared__ uint32_t part[kReduceLanes][kReduceDwords];
    const uint64_t c = blockIdx.z;
    const uint32_t row = blockIdx.y;
    const uint32_t dw = threadIdx.x % kReduceDwords, lane = threadIdx.x / kReduceDwords;
    const uint64_t d = static_cast<uint64_t>(blockIdx.x) * kReduceDwords + dw;
    const uint64_t words = a.L / 4;
    const bool dec = a.dst_ptrs != nullptr;
    if (dec && row >= load_uniform(a.rows + c)) return;  // uniform over the workgroup
    uint32_t v = 0;
    if (d < words) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.partials) + (c * a.partial_rows + row) * a.segs * words + d;
#pragma unroll 8
        for (uint32_t sg = lane; sg < a.segs; sg += kReduceLanes) v ^= *gbl<uint32_t>(src + sg * words);
    }
    part[lane][dw] = v;
    __syncthreads();
    for (uint32_t h = kReduceLanes / 2; h > 0; h >>= 1) {
        if (lane < h) part[lane][dw] ^= part[lane + h][dw];
        __syncthreads();
    }
    if (lane != 0 || d >= words) return;
    uint8_t *dst = dec ? load_uniform(a.dst_ptrs + c * a.rmax + row)
                       : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(row) * a.L;
    *gbl_mut<uint32_t>(dst + 4 * d) = part[0][dw];
}

template <bool VEC, bool DEC, int FULL>
__global__ __launch_bounds__(kThreads) void rows_terms16_kernel(RowsTermsArgs a, uint32_t tiles, uint32_t units) {
    __shared__ __align__(16) uint8_t lds[kTermChunk * 64 * 32];
    // XCD-aware: all tiles of one (codeword, row group) on one XCD.
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
    const uint64_t p = static_cast<uint64_t>(tile) * kTile + threadIdx.x * 16;
    if (DEC && rows_here == 0)
        copy_only16<VEC>(a, c, p, tb, te);
    else if (rows_here <= 4)
        pass16<VEC, DEC, 4, FULL>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
    else if (rows_here <= 8)
        pass16<VEC, DEC, 8, FULL>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
    else
        pass16<VEC, DEC, 16, FULL>(a, lds, c, rg, rows_here, p, do_copy, tb, te);
}

}  // namespace

hipError_t launch_reduce_partials(const RowsTermsArgs &a, hipStream_t s) {
    const uint64_t words = a.L / 4;
    if (a.segs >= 256 && a.count * a.partial_rows * words < 32768) {
        const dim3 rgrid(static_cast<uint32_t>((words + kReduceDwords - 1) / kReduceDwords), a.partial_rows,
                         static_cast<uint32_t>(a.count));
        hipLaunchKernelGGL(reduce_partials_wide_kernel, rgrid, dim3(256), 0, s, a);
    } else {
        const dim3 rgrid(static_cast<uint32_t>((words + 255) / 256), a.partial_rows, static_cast<uint32_t>(a.count));
        hipLaunchKernelGGL(reduce_partials_kernel, rgrid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_rows_terms16(const RowsTermsArgs &args, bool vec16, hipStream_t s) {
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
    const bool dec = a.src_ptrs != nullptr;
    if (dec != (a.dst_ptrs != nullptr) || dec != (a.copy_ptrs != nullptr) || dec != (a.rows != nullptr))
        return hipErrorInvalidValue;
    if (dec && a.accumulate) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(blocks));
    const uint32_t t32 = static_cast<uint32_t>(tiles), u32 = static_cast<uint32_t>(units);
    // Partial rows when the caller's buffer holds them (rows bounded by
    // partial_rows, L a multiple of 4 as every split requires).
    if (!split || !partials_usable(a)) a.partials = nullptr;
    if (split && !a.partials) {
        hipError_t err = launch_zero_dst_rows(a, s);
        if (err != hipSuccess) return err;
    }
    // Term counts that are multiples of 16 or 8 take the unrolled kernels,
    // split launches too when every segment holds whole chunks.
    const int full = !vec16 ? 0
                     : (a.T % 16 == 0 && a.seg_terms % 16 == 0 ? 16
                        : (a.T % 8 == 0 && a.seg_terms % 8 == 0 ? 8 : 0));
    if (full == 16 && dec)
        hipLaunchKernelGGL((rows_terms16_kernel<true, true, 16>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (full == 16)
        hipLaunchKernelGGL((rows_terms16_kernel<true, false, 16>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (full == 8 && dec)
        hipLaunchKernelGGL((rows_terms16_kernel<true, true, 8>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (full == 8)
        hipLaunchKernelGGL((rows_terms16_kernel<true, false, 8>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (vec16 && dec)
        hipLaunchKernelGGL((rows_terms16_kernel<true, true, 0>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (vec16)
        hipLaunchKernelGGL((rows_terms16_kernel<true, false, 0>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else if (dec)
        hipLaunchKernelGGL((rows_terms16_kernel<false, true, 0>), grid, dim3(kThreads), 0, s, a, t32, u32);
    else
        hipLaunchKernelGGL((rows_terms16_kernel<false, false, 0>), grid, dim3(kThreads), 0, s, a, t32, u32);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || !a.partials) return err;
    return launch_reduce_partials(a, s);
}

}  // namespace scrs
