// solve_kernels.hip - per-codeword Cauchy solve and decode preparation.
//
// For the k received symbols of a codeword, the r missing data rows are
// recovered from the r received repairs through the inverse of the r x r
// Cauchy submatrix A[i][j] = 1 / (row_i + col_j) (row = repair coordinate,
// col = missing data coordinate), fused with the contributions of the k - r
// present data rows.  The closed form is the reference's rational-Lagrange
// factorisation (scrs/src/decoder/cauchy_inverse.rs:29-170,
// scrs/src/tower/decoder.rs:433-546):
//   R_i = prod_h (row_i + col_h) / prod_{l != i} (row_i + row_l)
//   C_j = prod_l (col_j + row_l) / prod_{h != j} (col_j + col_h)
//   A^-1[j][i] = R_i C_j / (col_j + row_i)
//   present z:  F[j][z] = C_j * Q(z) / (z + col_j),
//               Q(z) = prod_h (z + col_h) / prod_l (z + row_l)
// Every entry is independent, so the whole solve is data parallel.  The
// fused coefficients are then expanded into split-nibble product tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "device.hpp"
#include "field.hpp"
#include "gf_device.hpp"
#include "kernels.hpp"

namespace scrs {
namespace {

constexpr int kSolveThreads = 256;

__constant__ uint8_t c_exp8[512];
__constant__ uint8_t c_log8[256];

__device__ __forceinline__ void stage_gf8(uint8_t *exp, uint8_t *log) {
    for (int i = threadIdx.x; i < 512; i += blockDim.x) exp[i] = c_exp8[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) log[i] = c_log8[i];
}

// Block-wide exclusive prefix sum of one 0/1 flag per thread; returns this
// thread's offset and writes the block total to *total.
__device__ __forceinline__ uint32_t block_scan(uint32_t flag, uint32_t *buf, uint32_t *total) {
    buf[threadIdx.x] = flag;
    __syncthreads();
    for (uint32_t off = 1; off < kSolveThreads; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0u;
        __syncthreads();
        buf[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t incl = buf[threadIdx.x];
    *total = buf[kSolveThreads - 1];
    __syncthreads();
    return incl - flag;
}

template <bool GF16> struct SolveLog;
template <> struct SolveLog<false> {
    using type = LogGf256;
    __device__ static type make(const uint8_t *exp, const uint8_t *log, const CauchySolveArgs &) {
        return LogGf256{{exp, log}};
    }
};
template <> struct SolveLog<true> {
    using type = LogGf65536;
    __device__ static type make(const uint8_t *, const uint8_t *, const CauchySolveArgs &a) {
        return LogGf65536{a.gf16_log_exp};
    }
};

// Codewords of more than this many symbols split the solve: the kernel
// below finds the rows, columns and their factors with one workgroup per
// codeword, and cauchy_fused_kernel computes the k x r fused entries with
// one thread per received position (one workgroup would serialise them).
constexpr uint32_t kSolveSplitK = 2048;

// The work buffer of codeword c: rpos (m u16, padded to even), then rv, cv,
// rf, cf (rmax u32 each), then q (k u32).
struct SolveWork {
    uint16_t *rpos;
    uint32_t *rv, *cv, *rf, *cf, *q;
    __device__ SolveWork(const CauchySolveArgs &a, uint64_t c) {
        rpos = a.work + c * a.work_stride;
        rv = reinterpret_cast<uint32_t *>(rpos + ((a.m + 1) & ~1u));
        cv = rv + a.rmax;
        rf = cv + a.rmax;
        cf = rf + a.rmax;
        q = cf + a.rmax;
    }
};

template <bool GF16, bool SPLIT>
__global__ __launch_bounds__(kSolveThreads) void cauchy_solve_kernel(CauchySolveArgs a) {
    __shared__ uint8_t exp[512];
    __shared__ uint8_t log[256];
    __shared__ uint32_t scan[kSolveThreads];
    __shared__ uint32_t s_r, s_e;
    extern __shared__ uint32_t seen[];  // n bits: index received
    using LF = typename SolveLog<GF16>::type;
    const LF L = SolveLog<GF16>::make(exp, log, a);
    const uint64_t c = blockIdx.x;
    const uint32_t k = a.k, n = a.k + a.m, r_cap = a.rmax;
    const uint16_t *idx = a.indices + c * k;
    stage_gf8(exp, log);
    const uint32_t words = (n + 31) / 32;
    for (uint32_t w = threadIdx.x; w < words; w += kSolveThreads) seen[w] = 0;
    __syncthreads();
    // The only pass over the indices.  Aligned index rows are read 8 at a
    // time, and bits of one bitmap word are merged before the LDS atomic
    // (received indices often come in runs).
    if ((reinterpret_cast<uintptr_t>(idx) & 15u) == 0 && k % 8 == 0) {
        const uint4 *v = reinterpret_cast<const uint4 *>(idx);
#pragma unroll 4
        for (uint32_t q8 = threadIdx.x; q8 < k / 8; q8 += kSolveThreads) {
            const uint4 x = v[q8];
            const uint32_t h[4] = {x.x, x.y, x.z, x.w};
            uint32_t word = (h[0] & 0xffffu) >> 5, bits = 0;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint32_t i = (h[e / 2] >> (16 * (e % 2))) & 0xffffu;
                if ((i >> 5) != word) {
                    atomicOr(&seen[word], bits);
                    word = i >> 5;
                    bits = 0;
                }
                bits |= 1u << (i & 31);
            }
            atomicOr(&seen[word], bits);
        }
    } else {
#pragma unroll 8
        for (uint32_t t = threadIdx.x; t < k; t += kSolveThreads) {
            const uint32_t i = idx[t];
            atomicOr(&seen[i >> 5], 1u << (i & 31));
        }
    }
    __syncthreads();
    const SolveWork wk(a, c);
    uint16_t *rpos = wk.rpos;  // [m] received repair k + x -> its row
    uint32_t *rv = wk.rv;      // [rmax] row (repair) coordinates
    uint32_t *cv = wk.cv;      // [rmax] column (missing data) coordinates
    uint32_t *rf = wk.rf;      // [rmax] row factors R_i
    uint32_t *cf = wk.cf;      // [rmax] column factors C_j
    uint32_t *q = wk.q;        // [k] Q(z) for present data, by received position
    uint16_t *missing = a.missing + c * r_cap;
    // Rows are the received repairs and columns the missing data indices,
    // both ascending: ranks are prefix counts over the received bitmap, one
    // run of consecutive words per thread.  (The order of the rows does not
    // change any fused entry: each is a function of the row and column sets
    // and of its own row.)
    const uint32_t wpt = (words + kSolveThreads - 1) / kSolveThreads;
    const uint32_t w0 = min(words, threadIdx.x * wpt), w1 = min(words, w0 + wpt);
    auto rep_bits = [&](uint32_t w) {  // received repairs in word w
        const uint32_t lo = w * 32, v = seen[w];
        const uint32_t below = lo >= k ? 0u : (k - lo >= 32 ? 0xffffffffu : (1u << (k - lo)) - 1u);
        const uint32_t valid = n - lo >= 32 ? 0xffffffffu : (1u << (n - lo)) - 1u;
        return v & ~below & valid;
    };
    auto miss_bits = [&](uint32_t w) {  // missing data indices in word w
        const uint32_t lo = w * 32;
        const uint32_t data = lo >= k ? 0u : (k - lo >= 32 ? 0xffffffffu : (1u << (k - lo)) - 1u);
        return ~seen[w] & data;
    };
    uint32_t nr = 0, nm = 0;
    for (uint32_t w = w0; w < w1; w++) {
        nr += __popc(rep_bits(w));
        nm += __popc(miss_bits(w));
    }
    uint32_t total_r, total_m;
    uint32_t row = block_scan(nr, scan, &total_r);
    uint32_t col = block_scan(nm, scan, &total_m);
    for (uint32_t w = w0; w < w1; w++) {
        for (uint32_t b = rep_bits(w); b; b &= b - 1) {
            const uint32_t i = w * 32 + __ffs(b) - 1;
            if (row < r_cap) {
                rpos[i - k] = static_cast<uint16_t>(row);
                rv[row] = a.coords[i];
            }
            row++;
        }
        for (uint32_t b = miss_bits(w); b; b &= b - 1) {
            const uint32_t i = w * 32 + __ffs(b) - 1;
            if (col < r_cap) {
                missing[col] = static_cast<uint16_t>(i);
                cv[col] = a.coords[i];
            }
            col++;
        }
    }
    if (threadIdx.x == 0) {
        s_r = total_r;
        s_e = total_m;
    }
    __syncthreads();
    const uint32_t r = s_r;
    if (r != s_e || r > r_cap) {
        if (threadIdx.x == 0) {
            atomicOr(a.error_flag, 1u);
            a.rows[c] = 0;  // later kernels then only copy the present rows
        }
        return;
    }
    // Log domain: rf, cf and q hold discrete logs; every product is a sum
    // of independent table lookups.  Coordinates are distinct, so every
    // difference is non-zero.
    constexpr uint32_t O = LF::kOrder;
    auto quot = [](uint32_t num, uint32_t den) { return (num % O + O - den % O) % O; };
    for (uint32_t i = threadIdx.x; i < r; i += kSolveThreads) {
        uint32_t cross = 0, within = 0;
        const uint32_t x = rv[i];
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) cross += L.lg(x ^ cv[h]);
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) within += l != i ? L.lg(x ^ rv[l]) : 0u;
        rf[i] = quot(cross, within);
    }
    for (uint32_t j = threadIdx.x; j < r; j += kSolveThreads) {
        uint32_t cross = 0, within = 0;
        const uint32_t x = cv[j];
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) cross += L.lg(x ^ rv[l]);
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) within += h != j ? L.lg(x ^ cv[h]) : 0u;
        cf[j] = quot(cross, within);
    }
    if (SPLIT) {
        if (threadIdx.x == 0) a.rows[c] = r;
        return;
    }
    for (uint32_t t = threadIdx.x; t < k; t += kSolveThreads) {
        const uint32_t i = idx[t];
        if (i >= k) continue;
        const uint32_t z = a.coords[i];
        uint32_t num = 0, den = 0;
#pragma unroll 8
        for (uint32_t h = 0; h < r; h++) num += L.lg(z ^ cv[h]);
#pragma unroll 8
        for (uint32_t l = 0; l < r; l++) den += L.lg(z ^ rv[l]);
        q[t] = quot(num, den);
    }
    __syncthreads();
    uint16_t *fused = a.fused + c * static_cast<uint64_t>(r_cap) * k;
#pragma unroll 4
    for (uint32_t x = threadIdx.x; x < r * k; x += kSolveThreads) {
        const uint32_t j = x / k, t = x % k;
        const uint32_t i = idx[t];
        uint32_t e, d;
        if (i >= k) {
            const uint32_t row = rpos[i - k];
            e = rf[row] + cf[j];
            d = L.lg(cv[j] ^ rv[row]);
        } else {
            e = cf[j] + q[t];
            d = L.lg(a.coords[i] ^ cv[j]);
        }
        e = e >= d ? e - d : e + O - d;  // < 2 * O
        if (GF16 && a.point_wlog) {
            // Weights w(indices[t]) / w(missing j), as logs.
            e = e % O + a.point_wlog[i] + (O - a.point_wlog[missing[j]]);  // < 3 * O
            e = e >= O ? e - O : e;
        }
        fused[j * k + t] = static_cast<uint16_t>(L.ex(e));
    }
    if (threadIdx.x == 0) a.rows[c] = r;
}

// Second half of a split solve: one thread per received position t of
// codeword blockIdx.y computes Q(z) (present data) or takes R_i (repair),
// then its fused entry for every missing row j (the same entries as the
// loop above).
template <bool GF16>
__global__ __launch_bounds__(kSolveThreads) void cauchy_fused_kernel(CauchySolveArgs a) {
    __shared__ uint8_t exp[512];
    __shared__ uint8_t log[256];
    using LF = typename SolveLog<GF16>::type;
    const LF L = SolveLog<GF16>::make(exp, log, a);
    if (!GF16) {
        stage_gf8(exp, log);
        __syncthreads();
    }
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k, t = blockIdx.x * kSolveThreads + threadIdx.x;
    const uint32_t r = a.rows[c];
    if (t >= k || r == 0) return;
    const SolveWork wk(a, c);
    const uint16_t *idx = a.indices + c * k;
    const uint16_t *missing = a.missing + c * a.rmax;
    uint16_t *fused = a.fused + c * static_cast<uint64_t>(a.rmax) * k;
    constexpr uint32_t O = LF::kOrder;
    const uint32_t i = idx[t];
    uint32_t base, z;
    if (i >= k) {
        const uint32_t row = wk.rpos[i - k];
        base = wk.rf[row];
        z = wk.rv[row];
    } else {
        z = a.coords[i];
        uint32_t num = 0, den = 0;
        for (uint32_t h = 0; h < r; h++) num += L.lg(z ^ wk.cv[h]);
        for (uint32_t l = 0; l < r; l++) den += L.lg(z ^ wk.rv[l]);
        base = (num % O + O - den % O) % O;
    }
    const uint32_t wi = (GF16 && a.point_wlog) ? a.point_wlog[i] : 0u;
    for (uint32_t j = 0; j < r; j++) {
        uint32_t e = base + wk.cf[j];  // < 2 * O
        const uint32_t d = L.lg(wk.cv[j] ^ z);
        e = e >= d ? e - d : e + O - d;  // < 2 * O
        if (GF16 && a.point_wlog) {
            e = e % O + wi + (O - a.point_wlog[missing[j]]);  // < 3 * O
            e = e >= O ? e - O : e;
        }
        fused[static_cast<uint64_t>(j) * k + t] = static_cast<uint16_t>(L.ex(e));
    }
}

// Split-nibble product tables for GF(2^8) from the fused coefficients:
// [c][t][g][n][16 rows] bytes, entry n < 16 = F * n, 16 + n = F * (n << 4).
// One thread per (t, g, row): eight basis products F * 2^b, then the 32
// entry bytes are XOR combinations of them.
__global__ __launch_bounds__(256) void tables_gf256_kernel(DecodePrepArgs a) {
    __shared__ uint8_t exp[512];
    __shared__ uint8_t log[256];
    stage_gf8(exp, log);
    __syncthreads();
    const Gf8Tables t8{exp, log};
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k, groups = a.groups;
    const uint32_t r = a.rows[c];
    const uint64_t items = static_cast<uint64_t>(k) * groups * 16;
    uint8_t *tab = reinterpret_cast<uint8_t *>(a.tables + c * static_cast<uint64_t>(k) * groups * 128);
    const uint16_t *fused = a.fused + c * static_cast<uint64_t>(a.rmax) * k;
    for (uint64_t x = blockIdx.x * 256ull + threadIdx.x; x < items; x += 256ull * gridDim.x) {
        const uint32_t s = x % 16, g = (x / 16) % groups, t = x / (16 * groups);
        const uint32_t row = g * 16 + s;
        if (g * 16 >= r) continue;  // the multiply-XOR never reads this group
        const uint32_t fv = row < r ? fused[row * k + t] : 0u;
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

// GF(2^16) tables: [c][t][g][q (4 nibbles)][n (16)][16 rows x 2 bytes].
// One thread per (t, g, row pair): dword stores of two rows' entries.  The
// products with the basis 2^b need no lookups: for b < 8, 2^b is the base
// field element x^b, which doubles both bytes of a + b u; for b >= 8 it is
// x^(b-8) u, and (a + b u) u = 0x20 b + (a + b) u (u^2 = u + 0x20).
__global__ __launch_bounds__(256) void tables_gf65536_kernel(DecodePrepArgs a) {
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k, groups = a.groups;
    const uint32_t r = a.rows[c];
    const uint64_t items = static_cast<uint64_t>(k) * groups * 8;
    uint32_t *tab = a.tables + c * static_cast<uint64_t>(k) * groups * 512;
    const uint16_t *fused = a.fused + c * static_cast<uint64_t>(a.rmax) * k;
    for (uint64_t x = blockIdx.x * 256ull + threadIdx.x; x < items; x += 256ull * gridDim.x) {
        const uint32_t s2 = x % 8, g = (x / 8) % groups, t = x / (8 * groups);
        if (g * 16 >= r) continue;  // the multiply-XOR never reads this group
        const uint32_t row = g * 16 + 2 * s2;
        const uint32_t f0 = row < r ? fused[row * k + t] : 0u;
        const uint32_t f1 = row + 1 < r ? fused[(row + 1) * k + t] : 0u;
        // Bytes of v: low and high byte of row 0, then of row 1.
        const uint32_t v = f0 | (f1 << 16);
        const uint32_t lo = v & 0x00ff00ffu, hi = (v >> 8) & 0x00ff00ffu;
        uint32_t hi20 = hi;
#pragma unroll
        for (int i = 0; i < 5; i++) hi20 = gf8_xtime4(hi20);  // 0x20 = x^5
        uint32_t basis[16];
        basis[0] = v;
        basis[8] = hi20 | ((lo ^ hi) << 8);
#pragma unroll
        for (int b = 1; b < 8; b++) {
            basis[b] = gf8_xtime4(basis[b - 1]);
            basis[8 + b] = gf8_xtime4(basis[7 + b]);
        }
        uint32_t *e = tab + (static_cast<uint64_t>(t) * groups + g) * 512 + s2;
#pragma unroll
        for (int qd = 0; qd < 4; qd++)
#pragma unroll
            for (int nib = 0; nib < 16; nib++) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if ((nib >> b) & 1) v ^= basis[4 * qd + b];
                e[(qd * 16 + nib) * 8] = v;
            }
    }
}

// Pointer tables of the multiply-XOR: received payloads, copies of present
// data rows, destinations of the missing rows.
// Grid (position chunks, codewords): one thread per received position.
__global__ __launch_bounds__(256) void pointers_kernel(DecodePrepArgs a) {
    const uint64_t c = blockIdx.y;
    const uint32_t k = a.k, m = a.m;
    const uint64_t L = a.L;
    const uint16_t *idx = a.indices + c * k;
    uint8_t *out = a.out_base + c * k * L;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < k) {
        const uint32_t i = idx[t];
        const uint8_t *src;
        if (a.payload_ptrs) src = a.payload_ptrs[c * k + t];
        else if (i < k) src = a.data_base + (c * k + i) * L;
        else src = a.repair_base + (c * m + (i - k)) * L;
        a.src_ptrs[c * k + t] = src;
        a.copy_ptrs[c * k + t] = i < k ? out + static_cast<uint64_t>(i) * L : nullptr;
    }
    if (blockIdx.x != 0) return;
    const uint32_t r = a.rows[c];
    const uint16_t *missing = a.missing + c * a.rmax;
    for (uint32_t s = threadIdx.x; s < r; s += blockDim.x)
        a.dst_ptrs[c * a.rmax + s] = out + static_cast<uint64_t>(missing[s]) * L;
}

}  // namespace

void gf8_exp_log(uint8_t exp[512], uint8_t log[256]) {
    uint8_t v = 1;
    for (int i = 0; i < 256; i++) log[i] = 0;
    for (int i = 0; i < 255; i++) {
        exp[i] = exp[i + 255] = v;
        log[v] = static_cast<uint8_t>(i);
        v = static_cast<uint8_t>(v ^ ((v << 1) ^ ((v & 0x80) ? 0x1B : 0)));
    }
    exp[510] = exp[0];
    exp[511] = exp[1];
}

hipError_t gf16_log_exp_device(const uint16_t **out) {
    static std::mutex mu;
    static std::map<int, DevBuf> tables;
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(mu);
    DevBuf &buf = tables[dev];
    if (!buf.ptr) {
        std::vector<uint16_t> t(kGf16LogExpEntries, 0);
        uint16_t v = 1;
        for (uint32_t i = 0; i < 65535u; i++) {
            t[v] = static_cast<uint16_t>(i);
            t[65536u + i] = t[65536u + 65535u + i] = v;
            v = Gf16::mul(v, Gf16::kGenerator);
        }
        if ((err = buf.ensure(t.size() * sizeof(uint16_t))) != hipSuccess) return err;
        if ((err = hipMemcpy(buf.ptr, t.data(), t.size() * sizeof(uint16_t), hipMemcpyHostToDevice)) != hipSuccess) {
            buf.release();
            return err;
        }
    }
    *out = buf.as<uint16_t>();
    return hipSuccess;
}

// The GF(2^8) constant tables are uploaded once per device (each device has
// its own copy of the __constant__ symbols).
static hipError_t ensure_solve_tables() {
    static std::mutex mu;
    static std::set<int> ready;
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(mu);
    if (ready.count(dev)) return hipSuccess;
    uint8_t exp[512], log[256];
    gf8_exp_log(exp, log);
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_exp8), exp, sizeof(exp))) != hipSuccess) return err;
    if ((err = hipMemcpyToSymbol(HIP_SYMBOL(c_log8), log, sizeof(log))) != hipSuccess) return err;
    ready.insert(dev);
    return hipSuccess;
}

uint64_t cauchy_solve_work_elems(uint32_t k, uint32_t m, uint32_t rmax) {
    // rpos (m u16, padded to even) + rv, cv, rf, cf (rmax u32 each) + q (k u32), in u16 units.
    return ((m + 1ull) & ~1ull) + 2ull * (4ull * rmax + k);
}

hipError_t launch_cauchy_solve(const CauchySolveArgs &a, bool gf16, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipError_t err = ensure_solve_tables();
    if (err != hipSuccess) return err;
    const size_t shmem = ((a.k + a.m + 31) / 32) * 4;
    if (shmem > 48 * 1024 || a.count > 0x7fffffffull) return hipErrorInvalidValue;
    CauchySolveArgs b = a;
    if (gf16 && (err = gf16_log_exp_device(&b.gf16_log_exp)) != hipSuccess) return err;
    const dim3 grid(static_cast<uint32_t>(a.count));
    if (a.k > kSolveSplitK && a.count <= 65535) {
        const dim3 grid2((a.k + kSolveThreads - 1) / kSolveThreads, static_cast<uint32_t>(a.count));
        if (gf16) {
            hipLaunchKernelGGL((cauchy_solve_kernel<true, true>), grid, dim3(kSolveThreads), shmem, s, b);
            if ((err = hipGetLastError()) != hipSuccess) return err;
            hipLaunchKernelGGL(cauchy_fused_kernel<true>, grid2, dim3(kSolveThreads), 0, s, b);
        } else {
            hipLaunchKernelGGL((cauchy_solve_kernel<false, true>), grid, dim3(kSolveThreads), shmem, s, b);
            if ((err = hipGetLastError()) != hipSuccess) return err;
            hipLaunchKernelGGL(cauchy_fused_kernel<false>, grid2, dim3(kSolveThreads), 0, s, b);
        }
    } else if (gf16) {
        hipLaunchKernelGGL((cauchy_solve_kernel<true, false>), grid, dim3(kSolveThreads), shmem, s, b);
    } else {
        hipLaunchKernelGGL((cauchy_solve_kernel<false, false>), grid, dim3(kSolveThreads), shmem, s, b);
    }
    return hipGetLastError();
}

hipError_t launch_decode_prepare(const DecodePrepArgs &a, bool gf16, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipError_t err = ensure_solve_tables();
    if (err != hipSuccess) return err;
    // Threads per (term, group): 16 rows (GF(2^8)) or 8 row pairs (GF(2^16)).
    const uint64_t items = static_cast<uint64_t>(a.k) * a.groups * (gf16 ? 8 : 16);
    if (a.count > 65535) return hipErrorInvalidValue;
    // At least 64 workgroups per codeword, more when few codewords would
    // leave the device idle (one codeword of tens of thousands of terms).
    const uint64_t bx_cap = std::max<uint64_t>(64, 2048 / a.count);
    const uint32_t bx = static_cast<uint32_t>(std::min<uint64_t>((items + 255) / 256, bx_cap));
    const dim3 grid(bx, static_cast<uint32_t>(a.count));
    // Without fused coefficients the tables already exist (a cached
    // recipe): only the pointer tables are written.
    if (a.fused) {
        if (gf16)
            hipLaunchKernelGGL(tables_gf65536_kernel, grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(tables_gf256_kernel, grid, dim3(256), 0, s, a);
        if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    hipLaunchKernelGGL(pointers_kernel, dim3((a.k + 255) / 256, static_cast<uint32_t>(a.count)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace scrs
