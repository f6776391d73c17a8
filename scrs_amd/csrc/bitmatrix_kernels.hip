// bitmatrix_kernels.hip - GF(2^8) multiply-XOR as a GF(2) bit-matrix product
// on the matrix cores of gfx950, with int8 or e2m1 (fp4) operands.
//
// Over GF(2) a GF(2^8) code of T sources and R rows is an (8R) x (8T) 0/1
// matrix M acting on the bit vector of every byte column:
//   bit i of dst_r = XOR over (t, j) of M[(r, i)][(t, j)] * bit j of src_t,
//   M[(r, i)][(t, j)] = bit i of (C[t][r] * 2^j)     (polynomial basis, 0x11B)
// which is the same multiply-XOR as the table kernels (the reference's
// rows_terms family, scrs/src/simd/x86.rs:658) computed another way.
//
// The product runs on v_mfma_i32_16x16x64_i8 with int32 accumulation.  The
// parity of each dot product is read from one accumulator bit without
// unpacking the data bytes to single bits: the data operand keeps bit j of a
// byte in place (x & 2^j, one AND) and the matrix operand carries the weight
// 2^(7-j) (as int8, 2^7 is -128), so every non-zero product is +-128 and the
// accumulator is 128 * (number of ones) mod 256: its bit 7 is the parity.
//
// Operands of one MFMA (16 x 16 x 64):
//   A (matrix, 16 rows x 64 K): lane l holds row r = l % 16 of output bit
//     plane i; K = 16 bytes: sources t0 = 8ch + 2(l/16) and t0 + 1, bits 0..7
//     of each.
//   B (data, 64 K x 16 columns): lane l holds column l % 16 with the same K
//     bytes, built from the two sources' bytes of that column.
// A and B distribute K over lanes identically, so each A byte meets the B
// byte of the same (source, bit).  The result D (16 rows x 16 columns): lane
// l holds column l % 16, rows 4(l/16) .. +3.
//
// A wave covers 64 columns per step: lane (q = l % 16) reads one dword of each
// source at column 4q and feeds byte u of it into N-tile u (u = 0..3), so the
// 4 N-tiles x 8 bit planes make 32 accumulator tiles per step.
//
// The e2m1 variant (bitmatrix_fp4_rows_kernel, below) runs the same product
// on v_mfma_f32_16x16x128_f8f6f4 with 4-bit operands: K = 128 = 16 sources x
// 8 bits per instruction, twice the int8 instruction's K at the same issue
// cost.  Products are 0 or 1 and the fp32 sums are exact.
#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "rows_device.hpp"

namespace scrs {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kBmThreads = 256;          // 4 wave64
constexpr int kBmWaveCols = 64;          // columns per wave per step
constexpr int kBmStepCols = kBmThreads / 64 * kBmWaveCols;  // 256 per workgroup step
constexpr int kBmTile = 4096;            // columns per workgroup
constexpr int kBmSteps = kBmTile / kBmStepCols;
constexpr int kBmChunkV4 = 8 * 64;       // 16-byte operands per 8-source chunk (8 planes x 64 lanes)
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t ld_dword(const uint8_t *p) {
    return __builtin_nontemporal_load(rowsdev::gbl<uint32_t>(p));
}

// NCH 8-source chunks (T <= 8 * NCH).  DEC: decode addressing (pointer
// tables, verbatim copies of the present data rows by row group 0).
template <int NCH, bool DEC>
__global__ __launch_bounds__(kBmThreads, 2) void bitmatrix_rows_kernel(RowsTermsArgs a, uint32_t tiles, uint32_t units) {
    extern __shared__ v4i lds_a[];  // [NCH][8 planes][64 lanes]
    __shared__ const uint8_t *s_src[8 * NCH];
    __shared__ uint8_t *s_copy[8 * NCH];
    __shared__ uint8_t *s_dst[16];

    const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t unit = xcd + kXcds * (slot / tiles);
    if (unit >= units) return;
    const uint32_t tile = slot % tiles;
    const uint64_t c = unit / a.groups_launched;
    const uint32_t g = unit % a.groups_launched;
    const uint32_t R = DEC ? a.rows[c] : a.rmax;
    const uint32_t base = g * 16;
    const uint32_t rows_here = R > base ? min(16u, R - base) : 0u;
    const bool do_copy = DEC && g == 0;
    if (rows_here == 0 && !do_copy) return;
    const uint32_t T = a.T;

    const v4i *tab = reinterpret_cast<const v4i *>(a.tables + c * a.table_cw_stride) +
                     static_cast<uint64_t>(g) * NCH * kBmChunkV4;
    for (uint32_t i = threadIdx.x; i < NCH * kBmChunkV4; i += kBmThreads) lds_a[i] = tab[i];
    for (uint32_t t = threadIdx.x; t < 8 * NCH; t += kBmThreads) {
        if (DEC) {
            s_src[t] = t < T ? a.src_ptrs[c * T + t] : nullptr;
            s_copy[t] = (t < T && do_copy) ? a.copy_ptrs[c * T + t] : nullptr;
        } else {
            s_src[t] = t < T ? a.src_base + c * a.src_cw_stride + static_cast<uint64_t>(t) * a.L : nullptr;
            s_copy[t] = nullptr;
        }
    }
    if (threadIdx.x < 16) {
        const uint32_t row = base + threadIdx.x;
        uint8_t *d = nullptr;
        if (threadIdx.x < rows_here)
            d = DEC ? a.dst_ptrs[c * a.rmax + row] : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(row) * a.L;
        s_dst[threadIdx.x] = d;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const uint32_t q = lane % 16, b = lane / 16;
    const uint64_t L = a.L;
    const uint64_t col0 = static_cast<uint64_t>(tile) * kBmTile + wave * kBmWaveCols + 4 * q;

    // Sources of this lane in chunk ch: 8ch + 2b and 8ch + 2b + 1 (pointers
    // re-read from LDS each step to keep them out of registers).
    auto load_step = [&](uint64_t col, uint32_t (&x)[2 * NCH]) {
        const bool active = col < L;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            const uint8_t *p0 = s_src[8 * ch + 2 * b], *p1 = s_src[8 * ch + 2 * b + 1];
            x[2 * ch] = (active && p0) ? ld_dword(p0 + col) : 0u;
            x[2 * ch + 1] = (active && p1) ? ld_dword(p1 + col) : 0u;
        }
    };

    uint32_t xa[2 * NCH], xb[2 * NCH];
    load_step(col0, xa);
    for (int step = 0; step < kBmSteps; step++) {
        const uint64_t col = col0 + static_cast<uint64_t>(step) * kBmStepCols;
        if (col - 4 * q >= L) break;  // the wave's 64 columns all lie beyond the symbol
        const bool active = col < L;
        // Prefetch the next step while this one multiplies.
        if (step + 1 < kBmSteps) load_step(col + kBmStepCols, xb);
        if (DEC && do_copy && active) {
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                uint8_t *c0 = s_copy[8 * ch + 2 * b], *c1 = s_copy[8 * ch + 2 * b + 1];
                if (c0) *rowsdev::gbl_mut<uint32_t>(c0 + col) = xa[2 * ch];
                if (c1) *rowsdev::gbl_mut<uint32_t>(c1 + col) = xa[2 * ch + 1];
            }
        }
        if (rows_here) {
            v4i acc[4][8];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int i = 0; i < 8; i++) acc[u][i] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                v4i B[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const uint32_t sel = 0x01010101u * u;
                    const uint32_t b0 = __builtin_amdgcn_perm(xa[2 * ch], xa[2 * ch], sel);
                    const uint32_t b1 = __builtin_amdgcn_perm(xa[2 * ch + 1], xa[2 * ch + 1], sel);
                    B[u] = v4i{static_cast<int>(b0 & 0x08040201u), static_cast<int>(b0 & 0x80402010u),
                               static_cast<int>(b1 & 0x08040201u), static_cast<int>(b1 & 0x80402010u)};
                }
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const v4i A = lds_a[(ch * 8 + i) * 64 + lane];
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        acc[u][i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B[u], acc[u][i], 0, 0, 0);
                }
            }
            // Parity extraction: bit 7 of accumulator (u, i, v) is bit i of
            // output byte (row 4b + v, column col + u).
#pragma unroll
            for (int v = 0; v < 4; v++) {
                uint32_t o = 0;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const uint32_t g01 = __builtin_amdgcn_perm(static_cast<uint32_t>(acc[1][i][v]),
                                                               static_cast<uint32_t>(acc[0][i][v]), 0x0c0c0400u);
                    const uint32_t g23 = __builtin_amdgcn_perm(static_cast<uint32_t>(acc[3][i][v]),
                                                               static_cast<uint32_t>(acc[2][i][v]), 0x04000c0cu);
                    o |= ((g01 | g23) >> (7 - i)) & (0x01010101u << i);
                }
                uint8_t *dst = s_dst[4 * b + v];
                if (active && dst) *rowsdev::gbl_mut<uint32_t>(dst + col) = o;
            }
        }
#pragma unroll
        for (int i = 0; i < 2 * NCH; i++) xa[i] = xb[i];
    }
}

// Decode tables from the fused coefficients: codeword c, row group g, chunk
// ch, plane i, lane l, byte p = weight(j) if bit i of F[row][t] * 2^j is set,
// with row = 16g + l % 16, t = 8ch + 2(l/16) + p/8, j = p % 8.
__global__ __launch_bounds__(256) void bitmatrix_tables_kernel(DecodePrepArgs a, uint32_t nch) {
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k;
    const uint32_t r = a.rows[c];
    const uint64_t dwords = static_cast<uint64_t>(a.groups) * nch * kBmChunkV4 * 4;
    uint32_t *tab = a.tables + c * dwords;
    const uint16_t *fused = a.fused + c * static_cast<uint64_t>(a.rmax) * k;
    for (uint64_t x = blockIdx.x * 256ull + threadIdx.x; x < dwords; x += 256ull * gridDim.x) {
        const uint32_t dq = x % 4, lane = (x / 4) % 64, i = (x / 256) % 8;
        const uint32_t ch = (x / 2048) % nch, g = static_cast<uint32_t>(x / (2048ull * nch));
        const uint32_t row = 16 * g + lane % 16;
        const uint32_t t = 8 * ch + 2 * (lane / 16) + dq / 2;
        uint32_t f = (row < r && t < k) ? fused[row * k + t] : 0u;
        const uint32_t j0 = (dq & 1) * 4;
        for (uint32_t j = 0; j < j0; j++) f = ((f << 1) ^ ((f & 0x80) ? 0x1Bu : 0u)) & 0xffu;
        uint32_t w = 0;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            if ((f >> i) & 1) w |= (0x80u >> (j0 + jj)) << (8 * jj);
            f = ((f << 1) ^ ((f & 0x80) ? 0x1Bu : 0u)) & 0xffu;
        }
        tab[x] = w;
    }
}

// ---- e2m1 (fp4) operands --------------------------------------------------
//
// Data operand.  A source byte x becomes one dword of 8 nibbles, nibble n
// holding bit fp4_bit(n) of x in place: the dword is (x, x, x, x >> 1) as
// bytes, masked with 0x44442211.  As e2m1 a set bit there reads 0.5 (nibble
// bit 0), 1 (bit 1) or 2 (bit 2); nibble bit 3 would be the sign, hence the
// shifted fourth byte.  The matrix operand carries the reciprocal weights
// (fp4_code: 2, 1, 0.5), so every non-zero product is exactly 1.0 and
// an accumulator counts the ones of its dot product.
//
// Parity.  Accumulators start at 2^23 and plane i's matrix operand runs
// with the e8m0 scale 2^i (scale byte 127 + i; 127 is 1.0), so the fp32 value
// 2^23 + count * 2^i (exact: counts stay below 2^16) has the count's unit at
// mantissa bit i, and bit i of the raw float is bit i of the output byte.
// Eight planes merge with bit-field selects, four columns with byte perms.
//
// Operands of one MFMA (16 x 16 x 128, e2m1 A and B in 4 VGPRs each):
//   A: lane l holds row l % 16 of a bit plane, K elements 32 (l / 16) ..
//     +31: sources 16ch + 4 (l / 16) + d (dword d), bits in nibble order.
//   B: lane l holds column l % 16 with the same K elements.
//   D: lane l holds column l % 16, rows 4 (l / 16) .. +3.
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kFp4Src = 16;                    // sources per chunk (K = 128)
// Nibble n of an operand dword: source bit (n >> 1) + 4 (n & 1), i.e.
// 0, 4, 1, 5, 2, 6, 3, 7, and matrix weight code 2.0, 2.0, 1.0, 1.0, 0.5 ...
__host__ __device__ constexpr uint32_t fp4_bit(uint32_t n) { return (n >> 1) + 4 * (n & 1); }
__host__ __device__ constexpr uint32_t fp4_code(uint32_t n) { return n < 2 ? 4u : (n < 4 ? 2u : 1u); }

// Operand dword of byte u of each source dword x (column u of 4).
__device__ __forceinline__ uint32_t fp4_data(uint32_t x, uint32_t u) {
    // Selector bytes u, u, u (x) and 4 + u (x >> 1).
    return __builtin_amdgcn_perm(x >> 1, x, 0x00010101u * u + ((4u + u) << 24)) & 0x44442211u;
}

// W: source dwords per lane per step (columns 4W per lane, 64W per wave).
// The four N-tiles of a sub-step run in two halves of two, which bounds the
// accumulators at 64 registers; from 8 chunks on the prefetched sources need
// the whole register file of one wave per SIMD.
template <int NCH, int W, int RING, bool DEC>
__global__ __launch_bounds__(kBmThreads, NCH <= 7 ? 2 : 1) void bitmatrix_fp4_rows_kernel(RowsTermsArgs a, uint32_t tiles,
                                                                          uint32_t units) {
    extern __shared__ v4i lds_a[];  // [NCH][8 planes][64 lanes]
    __shared__ const uint8_t *s_src[kFp4Src * NCH];
    __shared__ uint8_t *s_copy[kFp4Src * NCH];
    __shared__ uint8_t *s_dst[16];

    const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t unit = xcd + kXcds * (slot / tiles);
    if (unit >= units) return;
    const uint32_t tile = slot % tiles;
    const uint64_t c = unit / a.groups_launched;
    const uint32_t g = unit % a.groups_launched;
    const uint32_t R = DEC ? a.rows[c] : a.rmax;
    const uint32_t base = g * 16;
    const uint32_t rows_here = R > base ? min(16u, R - base) : 0u;
    const bool do_copy = DEC && g == 0;
    if (rows_here == 0 && !do_copy) return;
    const uint32_t T = a.T;

    const v4i *tab = reinterpret_cast<const v4i *>(a.tables + c * a.table_cw_stride) +
                     static_cast<uint64_t>(g) * NCH * kBmChunkV4;
    for (uint32_t i = threadIdx.x; i < NCH * kBmChunkV4; i += kBmThreads) lds_a[i] = tab[i];
    for (uint32_t t = threadIdx.x; t < kFp4Src * NCH; t += kBmThreads) {
        if (DEC) {
            s_src[t] = t < T ? a.src_ptrs[c * T + t] : nullptr;
            s_copy[t] = (t < T && do_copy) ? a.copy_ptrs[c * T + t] : nullptr;
        } else {
            s_src[t] = t < T ? a.src_base + c * a.src_cw_stride + static_cast<uint64_t>(t) * a.L : nullptr;
            s_copy[t] = nullptr;
        }
    }
    if (threadIdx.x < 16) {
        const uint32_t row = base + threadIdx.x;
        uint8_t *d = nullptr;
        if (threadIdx.x < rows_here)
            d = DEC ? a.dst_ptrs[c * a.rmax + row] : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(row) * a.L;
        s_dst[threadIdx.x] = d;
    }
    __syncthreads();

    constexpr uint32_t kWaveCols = 64 * W;                       // 16 lanes x 4W bytes
    constexpr uint32_t kStepCols = kBmThreads / 64 * kWaveCols;  // four waves
    constexpr int kSteps = kBmTile / kStepCols;
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const uint32_t q = lane % 16, b = lane / 16;
    const uint64_t L = a.L;
    const uint64_t col0 = static_cast<uint64_t>(tile) * kBmTile + wave * kWaveCols + 4 * W * q;

    // This lane's sources: 16ch + 4b + d.
    auto load_step = [&](uint64_t col, uint32_t (&x)[NCH][4][W]) {
#pragma unroll
        for (int ch = 0; ch < NCH; ch++)
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const uint8_t *p = s_src[kFp4Src * ch + 4 * b + d];
#pragma unroll
                for (int k = 0; k < W; k++) x[ch][d][k] = (p && col + 4 * k < L) ? ld_dword(p + col + 4 * k) : 0u;
            }
    };

    // One step: copies of present data rows, then the product of this
    // lane's columns.
    auto run_step = [&](uint64_t col, const uint32_t (&x)[NCH][4][W]) {
        if (DEC && do_copy) {
#pragma unroll
            for (int ch = 0; ch < NCH; ch++)
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    uint8_t *cp = s_copy[kFp4Src * ch + 4 * b + d];
                    if (!cp) continue;
#pragma unroll
                    for (int k = 0; k < W; k++)
                        if (col + 4 * k < L) *rowsdev::gbl_mut<uint32_t>(cp + col + 4 * k) = x[ch][d][k];
                }
        }
        if (rows_here) {
            uint32_t out[4][W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                uint32_t r[4][4];  // [v][u]: byte 0 is output row 4b + v, column u
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    v4f acc[2][8];
#pragma unroll
                    for (int ch = 0; ch < NCH; ch++) {
                        v8i B[2];
#pragma unroll
                        for (int uu = 0; uu < 2; uu++) {
                            const uint32_t u = 2 * h + uu;
                            B[uu] = v8i{static_cast<int>(fp4_data(x[ch][0][k], u)), static_cast<int>(fp4_data(x[ch][1][k], u)),
                                        static_cast<int>(fp4_data(x[ch][2][k], u)), static_cast<int>(fp4_data(x[ch][3][k], u)),
                                        0, 0, 0, 0};
                        }
                        // The matrix operands are re-read for every half and
                        // sub-step: an opaque copy of the lane index keeps the
                        // compiler from holding all of them in registers.
                        uint32_t al = lane;
                        asm volatile("" : "+v"(al));
#pragma unroll
                        for (int i = 0; i < 8; i++) {
                            const v4i a4 = lds_a[(ch * 8 + i) * 64 + al];
                            const v8i A = {a4[0], a4[1], a4[2], a4[3], 0, 0, 0, 0};
#pragma unroll
                            for (int uu = 0; uu < 2; uu++)
                                acc[uu][i] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                                    A, B[uu], ch == 0 ? v4f{8388608.0f, 8388608.0f, 8388608.0f, 8388608.0f} : acc[uu][i],
                                    4, 4, 0, 127 + i, 0, 127);
                        }
                    }
                    // Bit i of plane i's raw float is bit i of the output
                    // byte: seven bit selects per byte, each one v_bitop3
                    // (t & m) | (x & ~m), truth table f(0xF0, 0xCC, 0xAA) =
                    // 0xE4 for the operands (t, x, m).
#pragma unroll
                    for (int v = 0; v < 4; v++)
#pragma unroll
                        for (int uu = 0; uu < 2; uu++) {
                            uint32_t t = __float_as_uint(acc[uu][0][v]);
#pragma unroll
                            for (int i = 1; i < 8; i++)
                                t = __builtin_amdgcn_bitop3_b32(t, __float_as_uint(acc[uu][i][v]), (1u << i) - 1u, 0xE4);
                            r[v][2 * h + uu] = t;
                        }
                }
#pragma unroll
                for (int v = 0; v < 4; v++)
                    out[v][k] = __builtin_amdgcn_perm(r[v][1], r[v][0], 0x0c0c0400u) |
                                __builtin_amdgcn_perm(r[v][3], r[v][2], 0x04000c0cu);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                uint8_t *dst = s_dst[4 * b + v];
                if (!dst) continue;
#pragma unroll
                for (int k = 0; k < W; k++)
                    if (col + 4 * k < L) *rowsdev::gbl_mut<uint32_t>(dst + col + 4 * k) = out[v][k];
            }
        }
    };

    // Sources stream through a ring of RING steps, RING - 1 loads ahead.
    uint32_t xr[RING][NCH][4][W];
#pragma unroll
    for (int j = 0; j + 1 < RING; j++) load_step(col0 + static_cast<uint64_t>(j) * kStepCols, xr[j]);
    for (int s0 = 0; s0 < kSteps; s0 += RING) {
#pragma unroll
        for (int j = 0; j < RING; j++) {
            const int step = s0 + j;
            const uint64_t col = col0 + static_cast<uint64_t>(step) * kStepCols;
            if (col - 4 * W * q >= L) return;  // the wave's columns all lie beyond the symbol
            if (step + RING - 1 < kSteps) load_step(col + (RING - 1) * kStepCols, xr[(j + RING - 1) % RING]);
            run_step(col, xr[j]);
        }
    }
}

// fp4 operand dword (row, plane i, sources t0 + 0..3 as dwords d) for the
// coefficient f of one source: nibble n = fp4_code(n) if bit i of
// f * 2^fp4_bit(n) is set.
__host__ __device__ inline uint32_t fp4_matrix_dword(uint32_t f, uint32_t i) {
    uint32_t pw[8];
    for (int j = 0; j < 8; j++) {
        pw[j] = f;
        f = ((f << 1) ^ ((f & 0x80) ? 0x1Bu : 0u)) & 0xffu;
    }
    uint32_t w = 0;
    for (int n = 0; n < 8; n++)
        if ((pw[fp4_bit(n)] >> i) & 1) w |= fp4_code(n) << (4 * n);
    return w;
}

// Decode operands from the fused coefficients, in the fp4 layout: dword x of
// codeword c is (row group g, chunk ch, plane i, lane, d) with row 16g +
// lane % 16 and source 16ch + 4 (lane / 16) + d.
__global__ __launch_bounds__(256) void bitmatrix_fp4_tables_kernel(DecodePrepArgs a, uint32_t nch) {
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k;
    const uint32_t r = a.rows[c];
    const uint64_t dwords = static_cast<uint64_t>(a.groups) * nch * kBmChunkV4 * 4;
    uint32_t *tab = a.tables + c * dwords;
    const uint16_t *fused = a.fused + c * static_cast<uint64_t>(a.rmax) * k;
    for (uint64_t x = blockIdx.x * 256ull + threadIdx.x; x < dwords; x += 256ull * gridDim.x) {
        const uint32_t d = x % 4, lane = (x / 4) % 64, i = (x / 256) % 8;
        const uint32_t ch = (x / 2048) % nch, g = static_cast<uint32_t>(x / (2048ull * nch));
        const uint32_t row = 16 * g + lane % 16;
        const uint32_t t = kFp4Src * ch + 4 * (lane / 16) + d;
        const uint32_t f = (row < r && t < k) ? fused[row * k + t] : 0u;
        tab[x] = fp4_matrix_dword(f, i);
    }
}

#ifndef SCRS_FP4_RING
#define SCRS_FP4_RING 4  // source steps in flight per wave up to 4 chunks (a power of two)
#endif

template <int NCH>
hipError_t launch_fp4_nch(const RowsTermsArgs &a, bool dec, dim3 grid, uint32_t tiles, uint32_t units, hipStream_t s) {
    constexpr int W = 1;  // 8-byte source loads (W = 2) spill at these register budgets
    constexpr int RING = NCH <= 4 ? SCRS_FP4_RING : 2;  // registers allow the deep ring up to 4 chunks
    const size_t shmem = NCH * kBmChunkV4 * sizeof(v4i);
    hipError_t err =
        dec ? lds_opt_in(reinterpret_cast<const void *>(&bitmatrix_fp4_rows_kernel<NCH, W, RING, true>), 128 * 1024)
            : lds_opt_in(reinterpret_cast<const void *>(&bitmatrix_fp4_rows_kernel<NCH, W, RING, false>), 128 * 1024);
    if (err != hipSuccess) return err;
    if (dec)
        hipLaunchKernelGGL((bitmatrix_fp4_rows_kernel<NCH, W, RING, true>), grid, dim3(kBmThreads), shmem, s, a, tiles,
                           units);
    else
        hipLaunchKernelGGL((bitmatrix_fp4_rows_kernel<NCH, W, RING, false>), grid, dim3(kBmThreads), shmem, s, a, tiles,
                           units);
    return hipGetLastError();
}

template <int NCH>
hipError_t launch_nch(const RowsTermsArgs &a, bool dec, dim3 grid, uint32_t tiles, uint32_t units, hipStream_t s) {
    const size_t shmem = NCH * kBmChunkV4 * sizeof(v4i);
    // Above 64 KiB of dynamic LDS needs an explicit opt-in.
    hipError_t err = dec ? lds_opt_in(reinterpret_cast<const void *>(&bitmatrix_rows_kernel<NCH, true>), 96 * 1024)
                         : lds_opt_in(reinterpret_cast<const void *>(&bitmatrix_rows_kernel<NCH, false>), 96 * 1024);
    if (err != hipSuccess) return err;
    if (dec)
        hipLaunchKernelGGL((bitmatrix_rows_kernel<NCH, true>), grid, dim3(kBmThreads), shmem, s, a, tiles, units);
    else
        hipLaunchKernelGGL((bitmatrix_rows_kernel<NCH, false>), grid, dim3(kBmThreads), shmem, s, a, tiles, units);
    return hipGetLastError();
}

}  // namespace

static uint32_t bitmatrix_chunk_sources(uint32_t kind) { return kind == kBitmatrixFp4 ? kFp4Src : 8u; }

uint32_t bitmatrix_max_terms(uint32_t kind) {
    return (kind == kBitmatrixFp4 ? kFp4MaxChunks : kBitmatrixMaxChunks) * bitmatrix_chunk_sources(kind);
}

uint64_t bitmatrix_table_bytes(uint32_t T, uint32_t groups, uint32_t kind) {
    const uint32_t cs = bitmatrix_chunk_sources(kind);
    return static_cast<uint64_t>(groups) * ((T + cs - 1) / cs) * kBmChunkV4 * 16;
}

void build_bitmatrix_tables(const uint8_t *coef, uint32_t T, uint32_t R, uint32_t groups, uint8_t *out, uint32_t kind) {
    if (kind == kBitmatrixFp4) {
        const uint32_t nch = (T + kFp4Src - 1) / kFp4Src;
        uint32_t *o = reinterpret_cast<uint32_t *>(out);
        for (uint32_t g = 0; g < groups; g++)
            for (uint32_t ch = 0; ch < nch; ch++)
                for (uint32_t i = 0; i < 8; i++)
                    for (uint32_t lane = 0; lane < 64; lane++)
                        for (uint32_t d = 0; d < 4; d++) {
                            const uint32_t row = 16 * g + lane % 16, t = kFp4Src * ch + 4 * (lane / 16) + d;
                            const uint32_t f = (row < R && t < T) ? coef[static_cast<uint64_t>(t) * R + row] : 0u;
                            *o++ = fp4_matrix_dword(f, i);
                        }
        return;
    }
    const uint32_t nch = (T + 7) / 8;
    auto xt = [](uint32_t v) { return ((v << 1) ^ ((v & 0x80) ? 0x1Bu : 0u)) & 0xffu; };
    for (uint32_t g = 0; g < groups; g++)
        for (uint32_t ch = 0; ch < nch; ch++)
            for (uint32_t i = 0; i < 8; i++)
                for (uint32_t lane = 0; lane < 64; lane++)
                    for (uint32_t p = 0; p < 16; p++) {
                        const uint32_t row = 16 * g + lane % 16;
                        const uint32_t t = 8 * ch + 2 * (lane / 16) + p / 8, j = p % 8;
                        uint32_t f = (row < R && t < T) ? coef[static_cast<uint64_t>(t) * R + row] : 0u;
                        for (uint32_t s = 0; s < j; s++) f = xt(f);
                        const uint64_t off = ((((static_cast<uint64_t>(g) * nch + ch) * 8 + i) * 64 + lane) * 16) + p;
                        out[off] = ((f >> i) & 1) ? static_cast<uint8_t>(0x80u >> j) : 0;
                    }
}

hipError_t launch_rows_bitmatrix(const RowsTermsArgs &args, hipStream_t s, uint32_t kind) {
    if (args.count == 0 || args.L == 0) return hipSuccess;
    RowsTermsArgs a = args;
    if (a.L % 4 || a.T == 0 || a.T > bitmatrix_max_terms(kind) || a.accumulate || a.term_index)
        return hipErrorInvalidValue;
    const uint64_t tiles = (a.L + kBmTile - 1) / kBmTile;
    const uint32_t groups = a.rmax ? (a.rmax + 15) / 16 : 1;
    if (groups > a.groups + (a.rmax == 0)) return hipErrorInvalidValue;
    a.groups_launched = groups;
    const uint64_t units = a.count * groups;
    const uint64_t blocks = (units + kXcds - 1) / kXcds * kXcds * tiles;
    if (blocks > 0x7fffffffull || units > 0xffffffffull) return hipErrorInvalidValue;
    const bool dec = a.src_ptrs != nullptr;
    if (dec != (a.dst_ptrs != nullptr) || dec != (a.copy_ptrs != nullptr) || dec != (a.rows != nullptr))
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(blocks));
    const uint32_t t32 = static_cast<uint32_t>(tiles), u32 = static_cast<uint32_t>(units);
    if (kind == kBitmatrixFp4) {
        switch ((a.T + kFp4Src - 1) / kFp4Src) {
            case 1: return launch_fp4_nch<1>(a, dec, grid, t32, u32, s);
            case 2: return launch_fp4_nch<2>(a, dec, grid, t32, u32, s);
            case 3: return launch_fp4_nch<3>(a, dec, grid, t32, u32, s);
            case 4: return launch_fp4_nch<4>(a, dec, grid, t32, u32, s);
            case 5: return launch_fp4_nch<5>(a, dec, grid, t32, u32, s);
            case 6: return launch_fp4_nch<6>(a, dec, grid, t32, u32, s);
            case 7: return launch_fp4_nch<7>(a, dec, grid, t32, u32, s);
            case 8: return launch_fp4_nch<8>(a, dec, grid, t32, u32, s);
            case 9: return launch_fp4_nch<9>(a, dec, grid, t32, u32, s);
            case 10: return launch_fp4_nch<10>(a, dec, grid, t32, u32, s);
            case 11: return launch_fp4_nch<11>(a, dec, grid, t32, u32, s);
            case 12: return launch_fp4_nch<12>(a, dec, grid, t32, u32, s);
            case 13: return launch_fp4_nch<13>(a, dec, grid, t32, u32, s);
            case 14: return launch_fp4_nch<14>(a, dec, grid, t32, u32, s);
            case 15: return launch_fp4_nch<15>(a, dec, grid, t32, u32, s);
            default: return launch_fp4_nch<16>(a, dec, grid, t32, u32, s);
        }
    }
    switch ((a.T + 7) / 8) {
        case 1: return launch_nch<1>(a, dec, grid, t32, u32, s);
        case 2: return launch_nch<2>(a, dec, grid, t32, u32, s);
        case 3: return launch_nch<3>(a, dec, grid, t32, u32, s);
        case 4: return launch_nch<4>(a, dec, grid, t32, u32, s);
        case 5: return launch_nch<5>(a, dec, grid, t32, u32, s);
        case 6: return launch_nch<6>(a, dec, grid, t32, u32, s);
        case 7: return launch_nch<7>(a, dec, grid, t32, u32, s);
        default: return launch_nch<8>(a, dec, grid, t32, u32, s);
    }
}

hipError_t launch_bitmatrix_tables(const DecodePrepArgs &a, hipStream_t s, uint32_t kind) {
    if (a.count == 0) return hipSuccess;
    if (a.count > 65535) return hipErrorInvalidValue;
    const uint32_t cs = bitmatrix_chunk_sources(kind);
    const uint32_t nch = (a.k + cs - 1) / cs;
    const uint64_t dwords = static_cast<uint64_t>(a.groups) * nch * kBmChunkV4 * 4;
    const uint64_t want = (dwords + 255) / 256;
    const dim3 grid(static_cast<uint32_t>(want < 64 ? want : 64), static_cast<uint32_t>(a.count));
    if (kind == kBitmatrixFp4)
        hipLaunchKernelGGL(bitmatrix_fp4_tables_kernel, grid, dim3(256), 0, s, a, nch);
    else
        hipLaunchKernelGGL(bitmatrix_tables_kernel, grid, dim3(256), 0, s, a, nch);
    return hipGetLastError();
}

}  // namespace scrs
