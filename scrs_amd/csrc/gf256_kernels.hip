// gf256_kernels.hip - GF(2^8) multiply-XOR and decode-solve kernels for gfx950.
//
// The multiply-XOR computes, per byte column p of a codeword,
//   dst_r[p] = sum_t C[t][r] * src_t[p]            (GF(2^8), sum = XOR)
// which is the reference's rows_terms kernel family (scrs/src/simd/rows.rs:257,
// scrs/src/simd/x86.rs:658) applied to a whole batch of codewords at once.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace scrs {
namespace {

constexpr int kThreads = 256;                     // 4 wave64 per workgroup
constexpr int kBytesPerThread = 16;               // one 16-byte load per source
constexpr int kTile = kThreads * kBytesPerThread; // 4096 columns per workgroup
constexpr int kTermChunk = 32;                    // terms staged in LDS at once
constexpr int kRowsPerPass = 16;                  // 4 row quads per pass

// Product tables of one pass: [term][quad][lo 16 | hi 16] dwords = 16 KiB.
constexpr int kLdsDwords = kTermChunk * 4 * 32;

template <int NQ>
__device__ __forceinline__ void fold_term(uint32_t (&acc)[4][16], const uint32_t *tt,
                                          const uint32_t (&w)[4]) {
#pragma unroll
    for (int b = 0; b < 16; b++) {
        const uint32_t byte = (w[b >> 2] >> ((b & 3) * 8)) & 0xffu;
        const uint32_t lo = byte & 15u, hi = byte >> 4;
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[q][b] ^= tt[q * 32 + lo] ^ tt[q * 32 + 16 + hi];
    }
}

template <bool VEC>
__device__ __forceinline__ void load16(const uint8_t *src, uint64_t p, uint64_t L,
                                       uint32_t (&w)[4]) {
    if (VEC) {
        const uint4 v = *reinterpret_cast<const uint4 *>(src + p);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t d = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint64_t q = p + j * 4 + b;
                if (q < L) d |= static_cast<uint32_t>(src[q]) << (8 * b);
            }
            w[j] = d;
        }
    }
}

template <bool VEC>
__device__ __forceinline__ void store16(uint8_t *dst, uint64_t p, uint64_t L,
                                        const uint32_t (&w)[4]) {
    if (VEC) {
        *reinterpret_cast<uint4 *>(dst + p) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint64_t q = p + j * 4 + b;
                if (q < L) dst[q] = static_cast<uint8_t>(w[j] >> (8 * b));
            }
    }
}

template <bool VEC, int NQ>
__device__ void rows_terms_pass(const RowsTermsArgs &a, uint32_t *lds, uint64_t c, int rg,
                                int rows_here, uint64_t p, bool do_copy) {
    uint32_t acc[4][16];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int b = 0; b < 16; b++) acc[q][b] = 0;

    const bool active = p < a.L;
    const uint32_t *tab_cw = a.tables + c * a.table_cw_stride;
    for (uint32_t t0 = 0; t0 < a.T; t0 += kTermChunk) {
        const uint32_t tc = min(static_cast<uint32_t>(kTermChunk), a.T - t0);
        if (NQ > 0) {
            __syncthreads();
            // Stage the product tables of this term chunk and row group.
            constexpr uint32_t per_term = (NQ > 0 ? NQ : 1) * 32;
            for (uint32_t i = threadIdx.x; i < tc * per_term; i += kThreads) {
                const uint32_t t = i / per_term, rem = i % per_term;
                const uint32_t q = rem >> 5, n = rem & 31;
                lds[(t * 4 + q) * 32 + n] = tab_cw[((t0 + t) * a.quads + rg * 4 + q) * 32 + n];
            }
            __syncthreads();
        }
        if (!active) continue;
#pragma unroll 2
        for (uint32_t t = 0; t < tc; t++) {
            const uint64_t gt = t0 + t;
            const uint8_t *src = a.src_ptrs ? a.src_ptrs[c * a.T + gt]
                                            : a.src_base + c * a.src_cw_stride + gt * a.L;
            uint32_t w[4];
            load16<VEC>(src, p, a.L, w);
            if (do_copy) {
                uint8_t *cp = a.copy_ptrs[c * a.T + gt];
                if (cp) store16<VEC>(cp, p, a.L, w);
            }
            fold_term<NQ>(acc, lds + t * 128, w);
        }
    }
    if (!active) return;
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int r = rg * kRowsPerPass + q * 4 + s;
            if (q * 4 + s >= rows_here) continue;
            uint8_t *dst = a.dst_ptrs ? a.dst_ptrs[c * a.rmax + r]
                                      : a.dst_base + c * a.dst_cw_stride + static_cast<uint64_t>(r) * a.L;
            uint32_t o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                o[j] = ((acc[q][4 * j + 0] >> (8 * s)) & 0xffu) |
                       (((acc[q][4 * j + 1] >> (8 * s)) & 0xffu) << 8) |
                       (((acc[q][4 * j + 2] >> (8 * s)) & 0xffu) << 16) |
                       (((acc[q][4 * j + 3] >> (8 * s)) & 0xffu) << 24);
            }
            store16<VEC>(dst, p, a.L, o);
        }
}

template <bool VEC>
__global__ __launch_bounds__(kThreads) void rows_terms_lds_kernel(RowsTermsArgs a) {
    __shared__ uint32_t lds[kLdsDwords];
    const uint64_t c = blockIdx.y;
    const int rg = blockIdx.z;
    const uint32_t R = a.rows ? a.rows[c] : a.rmax;
    const int rows_here = static_cast<int>(min(static_cast<uint32_t>(kRowsPerPass),
                                               R > static_cast<uint32_t>(rg * kRowsPerPass)
                                                   ? R - rg * kRowsPerPass : 0u));
    const bool do_copy = a.copy_ptrs != nullptr && rg == 0;
    if (rows_here == 0 && !do_copy) return;
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * kTile + threadIdx.x * kBytesPerThread;
    const int nq = (rows_here + 3) / 4;
    switch (nq) {
        case 0: rows_terms_pass<VEC, 0>(a, lds, c, rg, rows_here, p, do_copy); break;
        case 1: rows_terms_pass<VEC, 1>(a, lds, c, rg, rows_here, p, do_copy); break;
        case 2: rows_terms_pass<VEC, 2>(a, lds, c, rg, rows_here, p, do_copy); break;
        case 3: rows_terms_pass<VEC, 3>(a, lds, c, rg, rows_here, p, do_copy); break;
        default: rows_terms_pass<VEC, 4>(a, lds, c, rg, rows_here, p, do_copy); break;
    }
}

// ---------------------------------------------------------------------------
// Decode solve: one wave64 per codeword.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint8_t gmul(const uint8_t *exp, const uint8_t *log, uint8_t x, uint8_t y) {
    return (x && y) ? exp[log[x] + log[y]] : 0;
}

__device__ __forceinline__ uint8_t ginv(const uint8_t *exp, const uint8_t *log, uint8_t x) {
    return x ? exp[255 - log[x]] : 0;
}

// GF(2^8) exp (512 entries, doubled) and log tables.
__constant__ uint8_t c_exp[512];
__constant__ uint8_t c_log[256];
bool g_tables_ready = false;

__global__ __launch_bounds__(64) void decode_solve_gf256_kernel(DecodeSolveArgs a) {
    extern __shared__ uint8_t smem[];
    uint8_t *exp = smem;            // 512
    uint8_t *log = exp + 512;       // 256
    uint8_t *seen = log + 256;      // 256
    uint8_t *rc = seen + 256;       // 256 repair columns, received order
    uint8_t *ms = rc + 256;         // 256 missing data rows, ascending
    uint8_t *tp = ms + 256;         // 256 position of received symbol t among the repairs
    uint8_t *aug = tp + 256;        // rmax x (2 * rmax) augmented matrix
    __shared__ uint32_t s_e;
    __shared__ uint32_t s_piv;

    const uint64_t c = blockIdx.x;
    const uint32_t k = a.k, m = a.m;
    const int lane = threadIdx.x;
    for (int i = lane; i < 512; i += 64) exp[i] = c_exp[i];
    for (int i = lane; i < 256; i += 64) { log[i] = c_log[i]; seen[i] = 0; }
    __syncthreads();
    const uint16_t *idx = a.indices + c * k;
    if (lane == 0) {
        uint32_t e = 0;
        for (uint32_t t = 0; t < k; t++) {
            const uint32_t i = idx[t];
            seen[i] = 1;
            tp[t] = static_cast<uint8_t>(e);
            if (i >= k) rc[e++] = static_cast<uint8_t>(i - k);
        }
        uint32_t nm = 0;
        for (uint32_t i = 0; i < k; i++)
            if (!seen[i]) ms[nm++] = static_cast<uint8_t>(i);
        s_e = e;
    }
    __syncthreads();
    const uint32_t e = s_e;
    const uint32_t w2 = 2 * e;
    // aug[t][s] = C[ms[s]][rc[t]], aug[t][e + s] = identity.
    for (uint32_t x = lane; x < e * w2; x += 64) {
        const uint32_t t = x / w2, s = x % w2;
        aug[t * w2 + s] = s < e ? a.coeffs[ms[s] * m + rc[t]] : static_cast<uint8_t>(s - e == t);
    }
    __syncthreads();
    // Gauss-Jordan; any elimination order yields the unique inverse.
    for (uint32_t col = 0; col < e; col++) {
        if (lane == 0) {
            uint32_t piv = col;
            while (piv < e && aug[piv * w2 + col] == 0) piv++;
            s_piv = piv;
        }
        __syncthreads();
        const uint32_t piv = s_piv;
        if (piv == e) {
            if (lane == 0) atomicOr(a.error_flag, 1u);
            return;
        }
        if (piv != col)
            for (uint32_t x = lane; x < w2; x += 64) {
                const uint8_t tmp = aug[col * w2 + x];
                aug[col * w2 + x] = aug[piv * w2 + x];
                aug[piv * w2 + x] = tmp;
            }
        __syncthreads();
        const uint8_t sc = ginv(exp, log, aug[col * w2 + col]);
        __syncthreads();
        for (uint32_t x = lane; x < w2; x += 64) aug[col * w2 + x] = gmul(exp, log, aug[col * w2 + x], sc);
        __syncthreads();
        for (uint32_t x = lane; x < e * w2; x += 64) {
            const uint32_t r = x / w2, cc = x % w2;
            if (r == col || cc == col) continue;
            const uint8_t f = aug[r * w2 + col];
            aug[r * w2 + cc] ^= gmul(exp, log, f, aug[col * w2 + cc]);
        }
        __syncthreads();
        for (uint32_t r = lane; r < e; r += 64)
            if (r != col) aug[r * w2 + col] = 0;
        __syncthreads();
    }
    // Binv[s][t] = aug[s][e + t].  The fused coefficient of received symbol
    // t' for missing row s is Binv[s][pos] for a repair and
    // sum_u Binv[s][u] * C[i][rc[u]] for present data row i; each is
    // expanded straight into the packed product tables.
    uint32_t *tab = a.tables + c * static_cast<uint64_t>(k) * a.quads * 32;
    for (uint32_t x = lane; x < k * a.quads * 32; x += 64) {
        const uint32_t t = x / (a.quads * 32), rem = x % (a.quads * 32);
        const uint32_t q = rem >> 5, n = rem & 31;
        const uint32_t i = idx[t];
        const uint8_t nib = static_cast<uint8_t>(n < 16 ? n : (n - 16) << 4);
        uint32_t word = 0;
        for (uint32_t s4 = 0; s4 < 4; s4++) {
            const uint32_t s = q * 4 + s4;
            if (s >= e) break;
            uint8_t f;
            if (i >= k) {
                // Received repair: its position among the received repairs.
                f = aug[s * w2 + e + tp[t]];
            } else {
                f = 0;
                for (uint32_t u = 0; u < e; u++)
                    f ^= gmul(exp, log, aug[s * w2 + e + u], a.coeffs[i * m + rc[u]]);
            }
            word |= static_cast<uint32_t>(gmul(exp, log, f, nib)) << (8 * s4);
        }
        tab[x] = word;
    }
    const uint64_t L = a.L;
    uint8_t *out = a.out_base + c * k * L;
    for (uint32_t t = lane; t < k; t += 64) {
        const uint32_t i = idx[t];
        const uint8_t *src;
        if (a.payload_ptrs) src = a.payload_ptrs[c * k + t];
        else if (i < k) src = a.data_base + (c * k + i) * L;
        else src = a.repair_base + (c * m + (i - k)) * L;
        a.src_ptrs[c * k + t] = src;
        a.copy_ptrs[c * k + t] = i < k ? out + static_cast<uint64_t>(i) * L : nullptr;
    }
    for (uint32_t s = lane; s < e; s += 64) a.dst_ptrs[c * a.rmax + s] = out + static_cast<uint64_t>(ms[s]) * L;
    if (lane == 0) a.rows[c] = e;
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

hipError_t launch_rows_terms_lds(const RowsTermsArgs &a, bool vec16, hipStream_t s) {
    if (a.count == 0 || a.L == 0) return hipSuccess;
    const uint64_t tiles = (a.L + kTile - 1) / kTile;
    const uint32_t groups = a.rmax ? (a.rmax + kRowsPerPass - 1) / kRowsPerPass : 1;
    if (tiles > 0x7fffffffull || a.count > 65535 || groups > 65535) return hipErrorInvalidValue;
    dim3 grid(static_cast<uint32_t>(tiles), static_cast<uint32_t>(a.count), groups);
    if (vec16)
        hipLaunchKernelGGL(rows_terms_lds_kernel<true>, grid, dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(rows_terms_lds_kernel<false>, grid, dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

static hipError_t ensure_field_tables() {
    if (g_tables_ready) return hipSuccess;
    uint8_t exp[512], log[256] = {0};
    uint8_t v = 1;
    for (int i = 0; i < 255; i++) {
        exp[i] = exp[i + 255] = v;
        log[v] = static_cast<uint8_t>(i);
        v = static_cast<uint8_t>(v ^ ((v << 1) ^ ((v & 0x80) ? 0x1B : 0)));
    }
    exp[510] = exp[0];
    exp[511] = exp[1];
    hipError_t err = hipMemcpyToSymbol(HIP_SYMBOL(c_exp), exp, sizeof(exp));
    if (err != hipSuccess) return err;
    err = hipMemcpyToSymbol(HIP_SYMBOL(c_log), log, sizeof(log));
    if (err == hipSuccess) g_tables_ready = true;
    return err;
}

hipError_t launch_decode_solve_gf256(const DecodeSolveArgs &a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipError_t err = ensure_field_tables();
    if (err != hipSuccess) return err;
    const size_t shmem = 512 + 256 * 5 + static_cast<size_t>(a.rmax) * 2 * a.rmax;
    hipLaunchKernelGGL(decode_solve_gf256_kernel, dim3(static_cast<uint32_t>(a.count)), dim3(64), shmem, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_splitmix(uint8_t *dst, uint64_t len, uint64_t seed, uint64_t byte_offset,
                                hipStream_t s) {
    if (len == 0) return hipSuccess;
    if (byte_offset % 8) return hipErrorInvalidValue;
    const uint64_t words = (len + 7) / 8;
    const uint64_t want = (words + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(fill_splitmix_kernel, dim3(blocks), dim3(256), 0, s, dst, len, seed, byte_offset / 8);
    return hipGetLastError();
}

}  // namespace scrs
