// batch_kernels.hip - persistent solve and multiply-XOR over an uneven batch.
//
// The multiply-XOR of a unit is the table path of the single-geometry
// kernels (split-nibble product tables staged in LDS, 16-byte column loads),
// with every source, destination and copy row taken from the range's
// pointer tables and every size from its descriptor.  When a range's terms
// fit one LDS chunk, a unit covering several codewords stages the table
// once for all of them.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "gf_device.hpp"
#include "rows_device.hpp"

namespace scrs {
namespace {

using namespace rowsdev;

constexpr uint32_t kChunk8 = 32;    // GF(2^8) terms per LDS fill (16 KiB at 16 rows)
constexpr uint32_t kChunk16 = 16;   // GF(2^16) terms per LDS fill (32 KiB at 16 rows)
constexpr uint32_t kLds = 32768;

// ---- solve ----------------------------------------------------------------

// Closed-form Cauchy solve of one decode job and the expansion of its fused
// coefficients into product tables; the same factorisation as the
// single-geometry solve (decoder/cauchy_inverse.rs:29-170).
// In the log domain (rf, cf and q hold discrete logs): long products become
// sums of independent lookups.  Coordinates are distinct, so no difference
// is zero.
template <typename LF>
__device__ void solve_job(const BatchSolve &j, const LF &L) {
    constexpr uint32_t O = LF::kOrder;
    auto quot = [](uint32_t num, uint32_t den) { return (num % O + O - den % O) % O; };
    const uint32_t k = j.k, r = j.r, tid = threadIdx.x, nt = blockDim.x;
    uint32_t *rv = j.work, *cv = rv + r, *rf = cv + r, *cf = rf + r, *q = cf + r;
    for (uint32_t t = tid; t < k; t += nt) {
        const uint32_t s = j.slot[t];
        if (s != 0xFFFFu) rv[s] = j.coords[j.idx[t]];
    }
    for (uint32_t i = tid; i < r; i += nt) cv[i] = j.coords[j.missing[i]];
    __syncthreads();
    for (uint32_t i = tid; i < r; i += nt) {
        uint32_t cross = 0, within = 0;
        const uint32_t x = rv[i];
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) cross += L.lg(x ^ cv[h]);
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) within += l != i ? L.lg(x ^ rv[l]) : 0u;
        rf[i] = quot(cross, within);
        uint32_t c2 = 0, w2 = 0;
        const uint32_t y = cv[i];
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) c2 += L.lg(y ^ rv[l]);
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) w2 += h != i ? L.lg(y ^ cv[h]) : 0u;
        cf[i] = quot(c2, w2);
    }
    for (uint32_t t = tid; t < k; t += nt) {
        if (j.slot[t] != 0xFFFFu) continue;
        const uint32_t z = j.coords[j.idx[t]];
        uint32_t num = 0, den = 0;
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) num += L.lg(z ^ cv[h]);
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) den += L.lg(z ^ rv[l]);
        q[t] = quot(num, den);
    }
    __syncthreads();
#pragma unroll 4
    for (uint32_t x = tid; x < r * k; x += nt) {
        const uint32_t jj = x / k, t = x % k;
        const uint32_t s = j.slot[t];
        uint32_t e, d;
        if (s != 0xFFFFu) {
            e = rf[s] + cf[jj];
            d = L.lg(cv[jj] ^ rv[s]);
        } else {
            e = cf[jj] + q[t];
            d = L.lg(j.coords[j.idx[t]] ^ cv[jj]);
        }
        e = e >= d ? e - d : e + O - d;  // < 2 * O
        j.fused[jj * k + t] = static_cast<uint16_t>(L.ex(e));
    }
    __syncthreads();
}

// Coefficient of term t for output row `row` of a table job.
template <typename F>
__device__ __forceinline__ uint32_t table_coef(const BatchSolve &j, const F &f, uint32_t t, uint32_t row) {
    if (row >= j.r) return 0u;
    if (j.encode) return f.inv(static_cast<uint32_t>(j.coords[t]) ^ j.coords[j.k + row]);
    return j.fused[row * j.k + t];
}

__device__ void tables_job8(const BatchSolve &j, const Gf8Tables &t8) {
    const uint32_t k = j.k, groups = j.groups;
    const FieldGf256 f{t8};
    uint8_t *tab = reinterpret_cast<uint8_t *>(j.tables);
    for (uint32_t x = threadIdx.x; x < k * groups * 16; x += blockDim.x) {
        const uint32_t s = x % 16, g = (x / 16) % groups, t = x / (16 * groups);
        const uint32_t fv = table_coef(j, f, t, g * 16 + s);
        uint32_t basis[8];
#pragma unroll
        for (int b = 0; b < 8; b++) basis[b] = t8.mul(fv, 1u << b);
        uint8_t *e = tab + (static_cast<uint64_t>(t) * groups + g) * 512 + s;
#pragma unroll
        for (int nib = 0; nib < 16; nib++) {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int b = 0; b < 4; b++)
                if ((nib >> b) & 1) { lo ^= basis[b]; hi ^= basis[4 + b]; }
            e[nib * 16] = static_cast<uint8_t>(lo);
            e[(16 + nib) * 16] = static_cast<uint8_t>(hi);
        }
    }
}

__device__ void tables_job16(const BatchSolve &j, const FieldGf65536 &f) {
    const uint32_t k = j.k, groups = j.groups;
    uint8_t *tab = reinterpret_cast<uint8_t *>(j.tables);
    for (uint32_t x = threadIdx.x; x < k * groups * 16; x += blockDim.x) {
        const uint32_t s = x % 16, g = (x / 16) % groups, t = x / (16 * groups);
        const uint32_t fv = table_coef(j, f, t, g * 16 + s);
        uint32_t basis[16];
#pragma unroll
        for (int b = 0; b < 16; b++) basis[b] = f.mul(fv, 1u << b);
        uint8_t *e = tab + (static_cast<uint64_t>(t) * groups + g) * 2048 + 2 * s;
#pragma unroll
        for (int qd = 0; qd < 4; qd++)
#pragma unroll
            for (int nib = 0; nib < 16; nib++) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if ((nib >> b) & 1) v ^= basis[4 * qd + b];
                uint8_t *p = e + (qd * 16 + nib) * 32;
                p[0] = static_cast<uint8_t>(v);
                p[1] = static_cast<uint8_t>(v >> 8);
            }
    }
}

__global__ __launch_bounds__(256) void batch_solve_kernel(const BatchHeader *hdr, const uint8_t *gf8,
                                                         const uint16_t *gf16_log_exp) {
    __shared__ uint8_t exp[512];
    __shared__ uint8_t log[256];
    for (uint32_t i = threadIdx.x; i < 512; i += blockDim.x) exp[i] = gf8[i];
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) log[i] = gf8[512 + i];
    __syncthreads();
    const Gf8Tables t8{exp, log};
    const uint32_t n = hdr->nsolve;
    const BatchSolve *solves = hdr->solves;
    for (uint32_t jx = blockIdx.x; jx < n; jx += gridDim.x) {
        const BatchSolve j = solves[jx];
        if (j.gf16) {
            if (!j.encode) solve_job(j, LogGf65536{gf16_log_exp});
            tables_job16(j, FieldGf65536{t8});
        } else {
            if (!j.encode) solve_job(j, LogGf256{t8});
            tables_job8(j, t8);
        }
        __syncthreads();
    }
}

// ---- multiply-XOR ----------------------------------------------------------

// Row r of codeword c's row group rg: its destination row, or the unit's
// own partial row when the range is term-split.
__device__ __forceinline__ uint8_t *partial_row(const BatchRange &R, uint32_t c, uint32_t row, uint32_t seg) {
    return R.partials + ((static_cast<uint64_t>(c) * R.rows + row) * R.segs + seg) * R.part_stride;
}

// 16 bytes of an output row at column p < L: the destination row, or the
// segment's partial row (16-byte aligned, part_stride >= L rounded up to 16).
template <bool VEC>
__device__ __forceinline__ void store_out(const BatchRange &R, uint32_t c, uint32_t row, uint32_t seg, uint64_t p,
                                          const uint32_t (&o)[4]) {
    if (R.segs > 1) {
        *gbl_mut<v4u32>(partial_row(R, c, row, seg) + p) = v4u32{o[0], o[1], o[2], o[3]};
    } else {
        store16<VEC>(load_uniform(R.dst + c * R.rows + row), p, R.L, o);
    }
}

template <bool VEC, bool DEC, int RW, int TPU>
__device__ __forceinline__ void unit8(const BatchRange &R, uint32_t cw0, uint32_t ncw, uint32_t rg, uint32_t rows_here, uint64_t p,
                      uint32_t tb, uint32_t te, uint32_t seg, uint8_t *lds) {
    constexpr uint32_t D = RW / 4;
    const bool active = p < R.L;
    const bool single = te - tb <= kChunk8;
    const bool do_copy = DEC && rg == 0;
    for (uint32_t c = cw0; c < cw0 + ncw; c++) {
        uint32_t acc[16][D];
#pragma unroll
        for (int b = 0; b < 16; b++)
#pragma unroll
            for (uint32_t d = 0; d < D; d++) acc[b][d] = 0;
        for (uint32_t t0 = tb; t0 < te; t0 += kChunk8) {
            const uint32_t tc = min(kChunk8, te - t0);
            // Two terms ahead: the term after next loads while this one
            // folds; the first two are issued before the chunk's tables are
            // staged, so their latency overlaps the staging's.
            uint32_t wn[4], wm[4];
            if (active) {
                load16<VEC>(load_uniform(R.src + c * R.T + t0), p, R.L, wn);
                if (tc > 1) load16<VEC>(load_uniform(R.src + c * R.T + t0 + 1), p, R.L, wm);
            }
            if (!single || c == cw0) {
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < tc * 32 * D; i += TPU) {
                    const uint32_t t = i / (32 * D), rem = i % (32 * D);
                    const uint32_t n = rem / D, d = rem % D;
                    reinterpret_cast<uint32_t *>(lds)[i] =
                        *gbl<uint32_t>(R.tables + ((t0 + t) * R.groups + rg) * 128 + n * 4 + d);
                }
                __syncthreads();
            }
            if (!active) continue;
#pragma unroll 1
            for (uint32_t t = 0; t < tc; t++) {
                const uint32_t gt = c * R.T + t0 + t;
                const uint32_t w[4] = {wn[0], wn[1], wn[2], wn[3]};
#pragma unroll
                for (int i = 0; i < 4; i++) wn[i] = wm[i];
                if (t + 2 < tc) load16<VEC>(load_uniform(R.src + gt + 2), p, R.L, wm);
                if (DEC && do_copy) {
                    uint8_t *cp = load_uniform(R.copy + gt);
                    if (cp) store16<VEC>(cp, p, R.L, w);
                }
                fold_term<RW>(acc, lds + t * 32 * RW, w);
            }
        }
        if (!active) continue;
#pragma unroll
        for (int r = 0; r < RW; r++) {
            if (static_cast<uint32_t>(r) >= rows_here) continue;
            const int d = r >> 2, s = (r & 3) * 8;
            uint32_t o[4];
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
                o[jj] = ((acc[4 * jj + 0][d] >> s) & 0xffu) | (((acc[4 * jj + 1][d] >> s) & 0xffu) << 8) |
                        (((acc[4 * jj + 2][d] >> s) & 0xffu) << 16) | (((acc[4 * jj + 3][d] >> s) & 0xffu) << 24);
            store_out<VEC>(R, c, rg * 16 + r, seg, p, o);
        }
    }
}

template <bool VEC, bool DEC, int RW, int TPU>
__device__ __forceinline__ void unit16(const BatchRange &R, uint32_t cw0, uint32_t ncw, uint32_t rg, uint32_t rows_here,
                       uint64_t p, uint32_t tb, uint32_t te, uint32_t seg, uint8_t *lds) {
    constexpr int EB = 2 * RW;
    constexpr uint32_t D = EB / 4;
    const bool active = p < R.L;
    const bool single = te - tb <= kChunk16;
    const bool do_copy = DEC && rg == 0;
    for (uint32_t c = cw0; c < cw0 + ncw; c++) {
        uint32_t acc[8][D];
#pragma unroll
        for (int e = 0; e < 8; e++)
#pragma unroll
            for (uint32_t d = 0; d < D; d++) acc[e][d] = 0;
        for (uint32_t t0 = tb; t0 < te; t0 += kChunk16) {
            const uint32_t tc = min(kChunk16, te - t0);
            uint32_t wn[4], wm[4];  // two terms ahead, as in unit8
            if (active) {
                load16<VEC>(load_uniform(R.src + c * R.T + t0), p, R.L, wn);
                if (tc > 1) load16<VEC>(load_uniform(R.src + c * R.T + t0 + 1), p, R.L, wm);
            }
            if (!single || c == cw0) {
                __syncthreads();
                const uint8_t *tb = reinterpret_cast<const uint8_t *>(R.tables);
                for (uint32_t i = threadIdx.x; i < tc * 64 * D; i += TPU) {
                    const uint32_t t = i / (64 * D), rem = i % (64 * D);
                    const uint32_t e = rem / D, d = rem % D;
                    reinterpret_cast<uint32_t *>(lds)[t * 64 * D + lds16_index<EB>(e, d)] =
                        *gbl<uint32_t>(tb + (static_cast<uint64_t>(t0 + t) * R.groups + rg) * 2048 + e * 32 + d * 4);
                }
                __syncthreads();
            }
            if (!active) continue;
#pragma unroll 1
            for (uint32_t t = 0; t < tc; t++) {
                const uint32_t gt = c * R.T + t0 + t;
                const uint32_t w[4] = {wn[0], wn[1], wn[2], wn[3]};
#pragma unroll
                for (int i = 0; i < 4; i++) wn[i] = wm[i];
                if (t + 2 < tc) load16<VEC>(load_uniform(R.src + gt + 2), p, R.L, wm);
                if (DEC && do_copy) {
                    uint8_t *cp = load_uniform(R.copy + gt);
                    if (cp) store16<VEC>(cp, p, R.L, w);
                }
                fold_term16<EB>(acc, lds + t * 64 * EB, w);
            }
        }
        if (!active) continue;
#pragma unroll
        for (int r = 0; r < RW; r++) {
            if (static_cast<uint32_t>(r) >= rows_here) continue;
            const int d = r >> 1, sh = (r & 1) * 16;
            uint32_t o[4];
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
                o[jj] = ((acc[2 * jj][d] >> sh) & 0xffffu) | (((acc[2 * jj + 1][d] >> sh) & 0xffffu) << 16);
            store_out<VEC>(R, c, rg * 16 + r, seg, p, o);
        }
    }
}

// Decode with no missing data row: only the present rows are copied.
template <bool VEC>
__device__ __forceinline__ void unit_copy(const BatchRange &R, uint32_t cw0, uint32_t ncw, uint64_t p, uint32_t tb, uint32_t te) {
    if (p >= R.L) return;
    for (uint32_t c = cw0; c < cw0 + ncw; c++)
        for (uint32_t t = tb; t < te; t++) {
            uint8_t *cp = load_uniform(R.copy + c * R.T + t);
            if (!cp) continue;
            uint32_t w[4];
            load16<VEC>(load_uniform(R.src + c * R.T + t), p, R.L, w);
            store16<VEC>(cp, p, R.L, w);
        }
}

template <bool VEC, bool DEC, int TPU>
__device__ __forceinline__ void run_unit(const BatchRange &R, uint32_t cw0, uint32_t ncw, uint32_t rg, uint32_t seg, uint64_t p,
                         uint8_t *lds) {
    const uint32_t base = rg * 16;
    const uint32_t rows_here = R.rows > base ? min(16u, R.rows - base) : 0u;
    const uint32_t tb = seg * R.seg_terms, te = min(R.T, tb + R.seg_terms);
    if (rows_here == 0) {
        if (DEC) unit_copy<VEC>(R, cw0, ncw, p, tb, te);
        return;
    }
    if (R.flags & kRangeGf16) {
        if (rows_here <= 4) unit16<VEC, DEC, 4, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
        else if (rows_here <= 8) unit16<VEC, DEC, 8, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
        else unit16<VEC, DEC, 16, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
    } else {
        if (rows_here <= 4) unit8<VEC, DEC, 4, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
        else if (rows_here <= 8) unit8<VEC, DEC, 8, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
        else unit8<VEC, DEC, 16, TPU>(R, cw0, ncw, rg, rows_here, p, tb, te, seg, lds);
    }
}

// A range read through the constant address space: scalar loads, so its
// fields stay in SGPRs for the whole unit.
__device__ __forceinline__ BatchRange load_range(const BatchRange *p) {
    static_assert(sizeof(BatchRange) % 4 == 0, "whole dwords");
    BatchRange r;
    const_u32_ptr s = (const_u32_ptr)(p);
    uint32_t *d = reinterpret_cast<uint32_t *>(&r);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(BatchRange) / 4; i++) d[i] = s[i];
    return r;
}

// Workgroup b runs unit b, then draws units from counter[0] (see
// launch_batch_rows): a workgroup without a first unit draws nothing, every
// other one draws until it is past the end.
template <int TPU>
__global__ __launch_bounds__(TPU) void batch_rows_kernel(const BatchHeader *hdr, uint32_t *counter) {
    __shared__ __align__(16) uint8_t lds[kLds];
    __shared__ uint32_t s_unit;
    const uint32_t nunits = hdr->nunits, nranges = hdr->nranges, base = hdr->claim_base;
    const BatchRange *ranges = hdr->ranges;
    uint32_t u = blockIdx.x;
    while (u < nunits) {
        // Last range whose first unit is <= u.
        uint32_t lo = 0, hi = nranges - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (load_uniform(&ranges[mid].unit_base) <= u) lo = mid;
            else hi = mid - 1;
        }
        // The search is wave-uniform; saying so keeps the range in SGPRs.
        const BatchRange R = load_range(ranges + __builtin_amdgcn_readfirstlane(lo));
        const uint32_t local = u - R.unit_base, per = R.groups * R.tiles * R.segs;
        const uint32_t chunk = local / per, rem = local % per;
        const uint32_t seg = rem % R.segs, rt = rem / R.segs;
        const uint32_t rg = rt / R.tiles, tile = rt % R.tiles;
        const uint32_t cw0 = chunk * R.ncw;
        const uint32_t ncw = min(R.ncw, R.cw_count - cw0);
        const uint64_t p = static_cast<uint64_t>(tile) * (TPU * 16) + threadIdx.x * 16;
        const bool vec = R.flags & kRangeVec16, dec = R.flags & kRangeDecode;
        if (vec && dec) run_unit<true, true, TPU>(R, cw0, ncw, rg, seg, p, lds);
        else if (vec) run_unit<true, false, TPU>(R, cw0, ncw, rg, seg, p, lds);
        else if (dec) run_unit<false, true, TPU>(R, cw0, ncw, rg, seg, p, lds);
        else run_unit<false, false, TPU>(R, cw0, ncw, rg, seg, p, lds);
        __syncthreads();
        // Draws past the first unit; the counter wraps, the difference does not.
        if (threadIdx.x == 0) s_unit = gridDim.x + (atomicAdd(counter, 1u) - base);
        __syncthreads();
        u = __builtin_amdgcn_readfirstlane(s_unit);
    }
}

// Reduce of the term-split ranges: item i is one dword of one destination
// row, the XOR of that dword over the row's segment partials.
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BatchHeader *hdr) {
    const uint32_t nred = hdr->nred;
    const uint64_t items = hdr->red_items;
    const BatchReduce *red = hdr->red;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < items;
         i += static_cast<uint64_t>(gridDim.x) * 256) {
        uint32_t lo = 0, hi = nred - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (red[mid].item_base <= i) lo = mid;
            else hi = mid - 1;
        }
        const BatchReduce &d = red[lo];
        const uint64_t words = d.L / 4, local = i - d.item_base;
        const uint64_t row = local / words, w = local - row * words;  // row = codeword * rows + row in codeword
        const uint8_t *src = d.partials + row * d.segs * d.stride + 4 * w;
        uint32_t v = 0;
#pragma unroll 4
        for (uint32_t sg = 0; sg < d.segs; sg++) v ^= *gbl<uint32_t>(src + sg * d.stride);
        *gbl_mut<uint32_t>(d.dst[row] + 4 * w) = v;
    }
}

}  // namespace

hipError_t launch_batch_solve(const BatchHeader *hdr, const uint8_t *gf8_tables, const uint16_t *gf16_log_exp,
                              uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(batch_solve_kernel, dim3(grid), dim3(256), 0, s, hdr, gf8_tables, gf16_log_exp);
    return hipGetLastError();
}

uint32_t batch_rows_grid(uint32_t tile_bytes) {
    // 256 CUs; LDS (32 KiB + a word) bounds residency at four workgroups.
    (void)tile_bytes;
    return 256 * 4;
}

hipError_t launch_batch_reduce(const BatchHeader *hdr, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(grid), dim3(256), 0, s, hdr);
    return hipGetLastError();
}

hipError_t launch_batch_rows(const BatchHeader *hdr, uint32_t *counter, uint32_t tile_bytes, uint32_t grid,
                             hipStream_t s) {
    switch (tile_bytes) {
        case 1024:
            hipLaunchKernelGGL(batch_rows_kernel<64>, dim3(grid), dim3(64), 0, s, hdr, counter);
            break;
        case 2048:
            hipLaunchKernelGGL(batch_rows_kernel<128>, dim3(grid), dim3(128), 0, s, hdr, counter);
            break;
        case 4096:
            hipLaunchKernelGGL(batch_rows_kernel<256>, dim3(grid), dim3(256), 0, s, hdr, counter);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace scrs
