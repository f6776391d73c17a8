// afft_kernels.hip - additive FFT over GF(2^16) for gfx950.
//
// Butterflies (scrs/src/afft/transform.rs:301-330, tower/payload.rs:58-130):
//   forward  lo' = lo + f*hi,  hi' = lo' + hi
//   inverse  hi' = lo + hi,    lo' = lo + f*hi'
// with f the factor of the butterfly's tree node.  Multiplication by a
// constant is GF(2)-linear, so c*x = sum over the four nibbles q of x of the
// table entry c*(n_q << 4q): a constant is a 64-entry u16 table (128 bytes,
// one dword per LDS bank, so a wave's lookups into one table never
// conflict).
//
// A pass owns the 2^g rows that differ in bits [b_lo, b_lo+g) of the row
// index, over a 256-byte column tile; each lane owns one dword column.  The
// g levels run in stages of up to 4 bits: a thread keeps the 16 rows of one
// stage tile in registers and tiles are exchanged through LDS between
// stages.  The node tables of the pass are built in LDS once, from the
// nodes' factors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>

#include "afft.hpp"
#include "device.hpp"
#include "afft_math.hpp"

namespace scrs {
namespace {

constexpr uint32_t kTileDwords = 64;  // 256-byte column tile
constexpr uint32_t kTab = 64;         // u16 entries per constant table
constexpr uint32_t kTabW = 128;       // u16 entries per wide constant table
// GF(2^16) as GF(2^8)[u] / (u^2 + u + 0x20), GF(2^8) modulo 0x11B; bit b
// of an element is x^b (b < 8) or u x^(b-8).
// v * x: both bytes times x in GF(2^8).
__device__ __forceinline__ uint32_t xtime16(uint32_t v) {
    return ((v << 1) & 0xFEFEu) ^ (((v >> 7) & 0x0101u) * 0x1Bu);
}
// v * u with u^2 = u + 0x20: (a + b u) u = 0x20 b + (a + b) u.
__device__ __forceinline__ uint32_t times_u(uint32_t v) {
    const uint32_t a = v & 0xffu, b = v >> 8;
    uint32_t lo = b;
#pragma unroll
    for (int i = 0; i < 5; i++) lo = xtime16(lo);
    return lo | ((a ^ b) << 8);
}
// Pass features, for kernels specialised on them.
enum : uint32_t {
    kFxInv = 1u << 0,      // inverse levels
    kFxFwd = 1u << 1,      // forward levels
    kFxDer = 1u << 2,      // derivative with its factors
    kFxUnit = 1u << 3,     // derivative with unit factors
    kFxSin = 1u << 4,      // input row scale
    kFxSout = 1u << 5,     // output row scale
    kFxSder = 1u << 6,     // derivative row scale
    kFxCopy = 1u << 7,     // copies of the loaded rows
    kFxAux = 1u << 8,      // derivative accumulator rows
    kFxInRows = 1u << 9,   // input row table
    kFxOutRows = 1u << 10, // output row table
    // Input row scale whose presence stays a run-time test: with it fixed,
    // the compiler schedules the decode's first pass into more registers
    // than 4 waves per SIMD allow, and its prefetched rows spill.
    kFxSinRt = 1u << 11,
    kFxDynamic = ~0u,      // features read from the arguments
};

typedef __attribute__((address_space(4))) const uint64_t *const_u64_ptr;
template <typename T>
__device__ __forceinline__ T *load_uniform(T *const *p) {
    return reinterpret_cast<T *>(*(const_u64_ptr)(p));
}

// Byte address of an LDS pointer, and a u16 read at a byte address.
__device__ __forceinline__ uint32_t lds_byte(const uint16_t *p) {
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>((const __attribute__((address_space(3))) uint16_t *)(p)));
}
__device__ __forceinline__ uint32_t lds_u16(uint32_t a) {
    return *(const __attribute__((address_space(3))) uint16_t *)(static_cast<uintptr_t>(a));
}

// Every constant table starts on a multiple of 128 bytes of the 128-byte
// aligned LDS block, so an entry's address is the table address OR the
// entry offset: one AND-OR per lookup after the shift.
__device__ __forceinline__ uint32_t gmul(const uint16_t *t, uint32_t x) {
    const uint32_t a = lds_byte(t);
    __builtin_assume((a & 127u) == 0);
    return lds_u16(a | ((x << 1) & 0x1Eu)) ^ lds_u16((a + 32u) | ((x >> 3) & 0x1Eu)) ^
           lds_u16((a + 64u) | ((x >> 7) & 0x1Eu)) ^ lds_u16((a + 96u) | ((x >> 11) & 0x1Eu));
}

// Two packed elements times the same constant.
__device__ __forceinline__ uint32_t gmul2(const uint16_t *t, uint32_t w) {
    return gmul(t, w & 0xffffu) | (gmul(t, w >> 16) << 16);
}

// Wide tables split an element 6 + 5 + 5 bits: entries [0, 64) are c*n,
// [64, 96) c*(n << 6), [96, 128) c*(n << 11).  A 64-entry part spans the 32
// banks once and a 32-entry part 16 of them, so lookups stay conflict-free
// with three reads per element instead of four.  Tables start on multiples
// of 256 bytes.
__device__ __forceinline__ uint32_t gmulw(const uint16_t *t, uint32_t x) {
    const uint32_t a = lds_byte(t);
    __builtin_assume((a & 255u) == 0);
    return lds_u16(a | ((x << 1) & 0x7Eu)) ^ lds_u16((a + 128u) | ((x >> 5) & 0x3Eu)) ^
           lds_u16((a + 192u) | ((x >> 10) & 0x3Eu));
}
__device__ __forceinline__ uint32_t gmul2w(const uint16_t *t, uint32_t w) {
    return gmulw(t, w & 0xffffu) | (gmulw(t, w >> 16) << 16);
}
template <bool W>
__device__ __forceinline__ uint32_t gmul2n(const uint16_t *t, uint32_t w) {
    if constexpr (W)
        return gmul2w(t, w);
    else
        return gmul2(t, w);
}

// The same products for the table at byte address `base + off`, where `base`
// is wave-uniform and aligned like a table and `off` is a compile-time
// multiple of the table size: the lookup address is (base | entry) + off,
// so `off` and the part offsets ride in the instruction's offset field and
// every table of a level shares one scalar base.
// The caller hides `base`'s value from the optimiser (an empty asm operand):
// knowing its alignment, the compiler would turn each OR into an add and
// move `off` and the part offsets back into scalar adds per table part.
template <bool W>
__device__ __forceinline__ uint32_t gmul1_at(uint32_t base, uint32_t off, uint32_t x) {
    if constexpr (W) {
        return lds_u16((base | ((x << 1) & 0x7Eu)) + off) ^ lds_u16((base | ((x >> 5) & 0x3Eu)) + off + 128u) ^
               lds_u16((base | ((x >> 10) & 0x3Eu)) + off + 192u);
    } else {
        return lds_u16((base | ((x << 1) & 0x1Eu)) + off) ^ lds_u16((base | ((x >> 3) & 0x1Eu)) + off + 32u) ^
               lds_u16((base | ((x >> 7) & 0x1Eu)) + off + 64u) ^ lds_u16((base | ((x >> 11) & 0x1Eu)) + off + 96u);
    }
}
template <bool W>
__device__ __forceinline__ uint32_t gmul2_at(uint32_t base, uint32_t off, uint32_t w) {
    return gmul1_at<W>(base, off, w & 0xffffu) | (gmul1_at<W>(base, off, w >> 16) << 16);
}

// gmul2_at split in two halves, so that a level's products can be software
// pipelined: the table reads of one product are issued before the previous
// product's reads are combined.
#ifndef SCRS_AFFT_PIPE
#define SCRS_AFFT_PIPE 1  // products in flight ahead of the one being combined
#endif
constexpr int kPipe = SCRS_AFFT_PIPE;

// f(integral_constant<int, I>) for I = 0 .. N-1, so that I is a constant
// expression inside f.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <bool W>
struct Parts {
    static constexpr int N = W ? 6 : 8;
    uint32_t r[N];
};
// Entry address of the field of `bits` bits at bit `lo` of w: one bit-field
// extract and one shift-OR into the table base.  Written out because the
// compiler rewrites the pair as shift + AND + OR, one instruction more
// where the mask is not an inline constant.  Two ALU instructions, no
// memory access.
template <uint32_t lo, uint32_t bits>
__device__ __forceinline__ uint32_t entry(uint32_t base, uint32_t w) {
    uint32_t a;
    asm("v_bfe_u32 %0, %1, %2, %3\n\tv_lshl_or_b32 %0, %0, 1, %4" : "=&v"(a) : "v"(w), "i"(lo), "i"(bits), "s"(base));
    return a;
}
template <bool W>
__device__ __forceinline__ void gmul2_issue(Parts<W> &p, uint32_t base, uint32_t off, uint32_t w) {
    const uint32_t x0 = w & 0xffffu, x1 = w >> 16;
    if constexpr (W) {
        p.r[0] = lds_u16(entry<0, 6>(base, w) + off);
        p.r[1] = lds_u16(entry<6, 5>(base, w) + off + 128u);
        p.r[2] = lds_u16(entry<11, 5>(base, w) + off + 192u);
        p.r[3] = lds_u16(entry<16, 6>(base, w) + off);
        p.r[4] = lds_u16(entry<22, 5>(base, w) + off + 128u);
        p.r[5] = lds_u16(entry<27, 5>(base, w) + off + 192u);
    } else {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t x = e ? x1 : x0;
            p.r[4 * e + 0] = lds_u16((base | ((x << 1) & 0x1Eu)) + off);
            p.r[4 * e + 1] = lds_u16((base | ((x >> 3) & 0x1Eu)) + off + 32u);
            p.r[4 * e + 2] = lds_u16((base | ((x >> 7) & 0x1Eu)) + off + 64u);
            p.r[4 * e + 3] = lds_u16((base | ((x >> 11) & 0x1Eu)) + off + 96u);
        }
    }
}
template <bool W>
__device__ __forceinline__ uint32_t gmul2_combine(const Parts<W> &p) {
    // The reads zero-extend, so the two 16-bit sums pack with one shift-OR:
    // two three-way XORs and the shift-OR (the compiler's 16-bit form takes
    // four instructions).
    if constexpr (W) {
        uint32_t lo, hi;
        asm("v_bitop3_b32 %0, %2, %3, %4 bitop3:0x96\n\t"
            "v_bitop3_b32 %1, %5, %6, %7 bitop3:0x96\n\t"
            "v_lshl_or_b32 %0, %1, 16, %0"
            : "=&v"(lo), "=&v"(hi)
            : "v"(p.r[0]), "v"(p.r[1]), "v"(p.r[2]), "v"(p.r[3]), "v"(p.r[4]), "v"(p.r[5]));
        return lo;
    }
    else
        return (p.r[0] ^ p.r[1] ^ p.r[2] ^ p.r[3]) | ((p.r[4] ^ p.r[5] ^ p.r[6] ^ p.r[7]) << 16);
}

// w AND m after taking w's upper half (SDWA operand select): the address
// field of the high element in one instruction.  Pure ALU.
__device__ __forceinline__ uint32_t and_hi(uint32_t m, uint32_t w) {
    uint32_t r;
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
        : "=v"(r)
        : "s"(m), "v"(w));
    return r;
}
// gmul2_issue (wide format) for a table at the compile-time LDS byte
// address `off`: with no base register to merge, the six entry addresses
// cost three shifts and six masks, the high element's through and_hi,
// instead of twelve instructions.
__device__ __forceinline__ void gmul2_issue_abs(Parts<true> &p, uint32_t off, uint32_t w) {
    const uint32_t w2 = w << 1, w5 = w >> 5, w10 = w >> 10;
    p.r[0] = lds_u16((w2 & 0x7Eu) + off);
    p.r[1] = lds_u16((w5 & 0x3Eu) + off + 128u);
    p.r[2] = lds_u16((w10 & 0x3Eu) + off + 192u);
    p.r[3] = lds_u16(and_hi(0x7Eu, w2) + off);
    p.r[4] = lds_u16(and_hi(0x3Eu, w5) + off + 128u);
    p.r[5] = lds_u16(and_hi(0x3Eu, w10) + off + 192u);
}
// Below the top stage a level's tables sit at a wave-uniform base.  A wide
// table's LDS byte address is a multiple of 256 and its entry offsets (with
// the field's table part in the instruction's offset field) are below 256,
// so an address register that holds the base keeps it in bits 8 and up,
// and each product rewrites only its low byte: an AND under SDWA with the
// other bytes preserved, one instruction per address as with compile-time
// addresses.  a.byte0 = w & m (lo) or w.hi & m (hi).  Pure ALU.
__device__ __forceinline__ void set_byte0(uint32_t &a, uint32_t m, uint32_t w) {
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:DWORD"
        : "+v"(a)
        : "s"(m), "v"(w));
}
__device__ __forceinline__ void set_byte0_hi(uint32_t &a, uint32_t m, uint32_t w) {
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:WORD_1"
        : "+v"(a)
        : "s"(m), "v"(w));
}
// The six address registers of a wide product (two elements, three fields).
struct WideAddr {
    uint32_t a[6];
};
__device__ __forceinline__ void wide_addr_init(WideAddr &ad, uint32_t base) {
#pragma unroll
    for (int f = 0; f < 6; f++) ad.a[f] = base;
}
// gmul2_issue (wide format) through the address registers of a table base
// (a multiple of 256); `off` is the table's constant offset from the base.
__device__ __forceinline__ void gmul2_issue_base(Parts<true> &p, WideAddr &ad, uint32_t off, uint32_t w) {
    const uint32_t w2 = w << 1, w5 = w >> 5, w10 = w >> 10;
    set_byte0(ad.a[0], 0x7Eu, w2);
    p.r[0] = lds_u16(ad.a[0] + off);
    set_byte0(ad.a[1], 0x3Eu, w5);
    p.r[1] = lds_u16(ad.a[1] + off + 128u);
    set_byte0(ad.a[2], 0x3Eu, w10);
    p.r[2] = lds_u16(ad.a[2] + off + 192u);
    set_byte0_hi(ad.a[3], 0x7Eu, w2);
    p.r[3] = lds_u16(ad.a[3] + off);
    set_byte0_hi(ad.a[4], 0x3Eu, w5);
    p.r[4] = lds_u16(ad.a[4] + off + 128u);
    set_byte0_hi(ad.a[5], 0x3Eu, w10);
    p.r[5] = lds_u16(ad.a[5] + off + 192u);
}
// Byte offset of level beta's first node table in a pass's block of node
// tables over G rows (tn u16 entries per table; mixed: the first level's
// tables in the nibble format).
constexpr uint32_t level_off_bytes(uint32_t G, uint32_t beta, bool mixed, uint32_t tn) {
    const uint32_t below = G - (G >> beta);
    return mixed ? (beta == 0 ? 0u : ((G / 2) * kTab + (below - G / 2) * kTabW) * 2u) : below * tn * 2u;
}

__device__ __forceinline__ uint32_t ld_col(const uint8_t *p, uint64_t cb, uint64_t L) {
    if (((reinterpret_cast<uintptr_t>(p) | L) & 3u) == 0) return *reinterpret_cast<const uint32_t *>(p + cb);
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (cb + b < L) w |= static_cast<uint32_t>(p[cb + b]) << (8 * b);
    return w;
}

__device__ __forceinline__ void st_col(uint8_t *p, uint64_t cb, uint64_t L, uint32_t w) {
    if (((reinterpret_cast<uintptr_t>(p) | L) & 3u) == 0) {
        *reinterpret_cast<uint32_t *>(p + cb) = w;
        return;
    }
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (cb + b < L) p[cb + b] = static_cast<uint8_t>(w >> (8 * b));
}

// A4 (every row address and L a multiple of 4, L < 2^31 so that a column
// offset past L still fits 32 bits): dword columns through buffer
// instructions.  A null row or a column at or past L is a zero load and a
// dropped store by the buffer's range, with no branch; the branchy byte
// paths above are kept for unaligned rows.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const uint8_t *row, const void *dummy, uint64_t L) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(row ? row : static_cast<const uint8_t *>(dummy)),
                                             static_cast<short>(0), row ? static_cast<uint32_t>(L) : 0u,
                                             0x00020000);
}
__device__ __forceinline__ uint32_t ld_col4(const uint8_t *row, const void *dummy, uint32_t cb, uint64_t L) {
    return __builtin_amdgcn_raw_buffer_load_b32(row_rsrc(row, dummy, L), cb, 0, 0);
}
__device__ __forceinline__ void st_col4(uint8_t *row, const void *dummy, uint32_t cb, uint64_t L, uint32_t w) {
    __builtin_amdgcn_raw_buffer_store_b32(w, row_rsrc(row, dummy, L), cb, 0, 0);
}

// Local row of register j in a stage over local bits [s_lo, s_lo + t) for
// wave `wv` holding R rows: tile u = wv*(R >> t) + (j >> t), i = j & (2^t-1).
template <int R>
__device__ __forceinline__ uint32_t rho(uint32_t wv, uint32_t s_lo, uint32_t t, uint32_t j) {
    const uint32_t u = wv * (static_cast<uint32_t>(R) >> t) + (j >> t);
    const uint32_t i = j & ((1u << t) - 1u);
    return (u & ((1u << s_lo) - 1u)) | (i << s_lo) | ((u >> s_lo) << (s_lo + t));
}

// One workgroup owns a row group and `tpb` consecutive column tiles: the
// node tables are built once and the tiles run in turn, the next tile's
// rows loading while the current one computes.
// GL: the pass's level count fixed at compile time (0: taken from the
// arguments), so a full pass's stage layouts and table offsets fold into
// constants.  W: node tables in the wide format.
#ifndef SCRS_AFFT_WAVES_PER_EU
#define SCRS_AFFT_WAVES_PER_EU 4
#endif
// A 16-row pass stages at least 64 KiB of LDS, so a CU holds one 1024-thread
// workgroup: 4 waves per SIMD.  Telling the compiler so lets it spend the
// registers of that occupancy on table reads in flight.
//
// MX (with W): mixed node tables, nibble tables for the pass's first level
// (half of the pass's nodes) and wide tables above it: three quarters of the
// wide footprint, so that a 7-level pass with inverse and forward tables
// still fits two workgroups per CU, with three reads per element on all but
// its first level.
// FX: the pass's features fixed at compile time (kFx* bits), or kFxDynamic
// to read them from the arguments.  A fixed set drops the branches of the
// absent features from the tile loop, and with them the register and
// wait-count bookkeeping the compiler keeps for every path: the flags and
// the absent pointers become constants, the present pointers non-null.
template <int R, bool A4, uint32_t GL, bool W, bool MX, uint32_t FX>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(SCRS_AFFT_WAVES_PER_EU, SCRS_AFFT_WAVES_PER_EU))) void afft_pass_kernel(
    AfftPassArgs a, uint32_t tpb) {
    if constexpr (FX != kFxDynamic) {
        a.flags = ((FX & kFxInv) ? kAfftInv : 0u) | ((FX & kFxFwd) ? kAfftFwd : 0u) |
                  ((FX & kFxDer) ? kAfftDeriv : 0u) | ((FX & kFxUnit) ? kAfftDerivUnit : 0u);
        auto fix = [](auto &ptr, bool present) {
            if (present)
                __builtin_assume(ptr != nullptr);
            else
                ptr = nullptr;
        };
        if (!(FX & kFxSinRt)) fix(a.scale_in, FX & kFxSin);
        fix(a.scale_out, FX & kFxSout);
        fix(a.scale_der, FX & kFxSder);
        fix(a.copy_rows, FX & kFxCopy);
        fix(a.aux_base, FX & kFxAux);
        fix(a.in_rows, FX & kFxInRows);
        fix(a.out_rows, FX & kFxOutRows);
    }
    constexpr uint32_t RB = R >= 16 ? 4 : (R >= 8 ? 3 : (R >= 4 ? 2 : (R >= 2 ? 1 : 0)));
    constexpr uint32_t kTN = W ? kTabW : kTab;          // u16 entries per node table
    static_assert(!MX || W, "mixed tables extend the wide format");
    extern __shared__ __align__(256) uint32_t smem[];  // tables at multiples of 128 (wide: 256) bytes
    const uint32_t g = GL ? GL : a.g, G = 1u << g, b_lo = a.b_lo;
    const uint32_t gidx = blockIdx.y;
    const uint32_t base = (gidx & ((1u << b_lo) - 1u)) | ((gidx >> b_lo) << (b_lo + g));
    if (a.row_limit && (base >> b_lo) > ((a.row_limit - 1u) >> b_lo)) return;
    const bool inv = a.flags & kAfftInv, fwd = a.flags & kAfftFwd;
    const bool unit = a.flags & kAfftDerivUnit, der = unit || (a.flags & kAfftDeriv);
    // A forward pass whose rows from row_limit on are not needed downstream,
    // when row_limit is half the pass's row range (the group's rows with the
    // pass's top bit set): after the top level those rows form a transform
    // of their own, whose butterflies and stores are skipped.
    const bool half = fwd && a.row_limit != 0 && a.row_limit == (1u << (b_lo + g - 1u)) && (gidx >> b_lo) == 0u;
    // Node tables: level beta's G >> (beta + 1) tables of tab_n(beta) u16
    // entries from level_off(beta) on.
    auto tab_n = [&](uint32_t beta) -> uint32_t { return (MX && beta == 0) ? kTab : kTN; };
    auto level_off = [&](uint32_t beta) -> uint32_t {
        const uint32_t below = G - (G >> beta);  // tables of the levels below beta
        if constexpr (MX)
            return beta == 0 ? 0u : (G / 2) * kTab + (below - G / 2) * kTabW;
        else
            return below * kTN;
    };
    // LDS: the node tables first, so that with the pass's size and features
    // fixed at compile time their addresses are constants (kAbs below), then
    // the other tables, then the data tile.
    const uint32_t tabs_n = level_off(g);
    uint16_t *inv_t = reinterpret_cast<uint16_t *>(smem);
    uint16_t *fwd_t = inv_t + (inv ? tabs_n : 0);
    uint16_t *der_t = fwd_t + (fwd ? tabs_n : 0);
    uint16_t *sin_t = der_t + (der && !unit ? 16 * kTab : 0);
    uint16_t *sout_t = sin_t + (a.scale_in ? G * kTab : 0);
    uint16_t *sder_t = sout_t + (a.scale_out ? G * kTab : 0);
    uint32_t *data = reinterpret_cast<uint32_t *>(sder_t + (a.scale_der ? G * kTab : 0));
    // Top stages with compile-time table addresses: a pass of GL levels with
    // a fixed feature set and wide tables.  kTabsByte(inverse): the LDS byte
    // address of that direction's node tables.
    constexpr bool kAbs = W && GL != 0 && FX != kFxDynamic;
    // Address registers for the wide levels below the top stage (the fixed
    // kernels; the general ones keep their registers for their branches).
    constexpr bool kBaseRegs = kAbs;
    constexpr uint32_t kGc = GL ? 1u << GL : 0u;
    constexpr uint32_t kFwdByte = (FX & kFxInv) ? level_off_bytes(kGc, GL, MX, kTN) : 0u;
    const uint32_t tid = threadIdx.x, nthr = (G / R) * 64u;  // the launch's block size

    // Row and derivative table staging in 16-byte chunks, kStageIn chunks
    // per thread in flight at once, so that one wait covers them all: chunk
    // c of a staged block comes from src_chunk(c).
    constexpr uint32_t kStageIn = 4;
    auto stage_chunks = [&](uint16_t *dst, uint32_t chunks, auto src_chunk) {
        uint4 *d = reinterpret_cast<uint4 *>(dst);
        for (uint32_t c0 = 0; c0 < chunks; c0 += kStageIn * nthr) {
            // Past the end, a thread loads and stores the last chunk again
            // (the same value to the same place): no branch around the
            // loads, which the compiler would otherwise sink each into its
            // store's branch, one wait per chunk.
            uint4 val[kStageIn];
#pragma unroll
            for (uint32_t q = 0; q < kStageIn; q++) val[q] = *src_chunk(min(c0 + q * nthr + tid, chunks - 1u));
#pragma unroll
            for (uint32_t q = 0; q < kStageIn; q++) d[min(c0 + q * nthr + tid, chunks - 1u)] = val[q];
        }
    };
    // The node tables of this group and its row-scale tables, built in LDS
    // from their factors.  Node tables: level beta (pass bit b_lo+beta)
    // uses G >> (beta+1) consecutive nodes, all levels as one block per
    // direction; table n of the block (in the pass's numbering, level by
    // level) is of level g - 1 - floor(log2(G - 1 - n)).  Row tables: one
    // per row of the group and scale present (input, output, derivative).
    // A table holds the products of its factor c with every value of each
    // address field, and c*v is the XOR of c*2^b over the bits b of v.
    //
    // A work item is a 32-entry quarter of a wide table (the halves of its
    // 6-bit field, its two 5-bit fields) or half of a nibble table (two
    // nibble fields).  Node items: 4 per table slot and direction, the
    // quarter outermost so that a wave's items are of one kind when a block
    // has 64 tables or more; a thread has at most two (8G items over
    // G / R * 64 >= 4G threads), items tid and tid + nthr.  Row items: 2 per
    // row, at most one per thread and scale.  The factors load before the
    // first tile's rows, so that the tables build while the rows arrive.
    const uint32_t lp = g + ((inv && fwd) ? 1u : 0u), per = 1u << lp;  // node items per quarter
    // Node of node item `item`, or ~0u for no table part.
    auto item_node = [&](uint32_t item) -> uint32_t {
        const uint32_t q = item >> lp, rest = item & (per - 1u), n = rest & (G - 1u);
        if (!(inv || fwd) || item >= 4u * per || n == G - 1u) return ~0u;
        const uint32_t beta = g - 1u - (31u - static_cast<uint32_t>(__builtin_clz(G - 1u - n)));
        if (tab_n(beta) != kTabW && q >= 2u) return ~0u;
        const uint32_t B = b_lo + beta, first = G - (G >> beta);
        const uint32_t pbase = ((!inv || rest >= G) ? a.fwd_base : a.inv_base) + base;
        return (1u << (15u - B)) + (pbase >> (B + 1)) + (n - first);
    };
    // Row of the group of row item tid: 2 items per row, the half outermost.
    const uint32_t r_item = base | ((tid & (G - 1u)) << b_lo);
    const bool row_item = tid < 2u * G;
    uint32_t fac[2], fac_r[3];
    auto load_factors = [&]() {
        // Every load unconditional, at a clamped address.
        if (inv || fwd) {
#pragma unroll
            for (uint32_t k = 0; k < 2; k++) {
                const uint32_t node = item_node(tid + k * nthr);
                fac[k] = a.node_fac[node == ~0u ? 0u : node];
            }
        }
        const uint32_t r = row_item ? r_item : base;
        if (a.scale_in) fac_r[0] = a.scale_in[r];
        if (a.scale_out) fac_r[1] = a.scale_out[r];
        if (a.scale_der) fac_r[2] = a.scale_der[r];
    };
    // The item of factor c: quarter (wide) or half (nibble) q, first entry
    // at dst.
    auto build_item = [&](uint32_t c, uint32_t q, bool wide, uint16_t *dst) {
        // Entry h*16 + m of the item (m < 16) is f[h] + sum over the bits b
        // of m of x[h][b].  Each kind of item forms only the c*2^b it uses
        // (a wave's items are of one kind when a block has 64 tables or more).
        uint32_t x[2][4], f[2];
        auto run = [&](uint32_t v, uint32_t(&o)[4]) {  // o[b] = v * 2^b
            o[0] = v;
#pragma unroll
            for (int b = 1; b < 4; b++) o[b] = xtime16(o[b - 1]);
        };
        if (wide && q < 2u) {  // bits 0..5
            run(c, x[0]);
            const uint32_t c16 = xtime16(x[0][3]), c32 = xtime16(c16);
            f[0] = q == 1u ? c32 : 0u;
            f[1] = f[0] ^ c16;
            run(c, x[1]);
        } else if (wide && q == 2u) {  // bits 6..10: c x^6, c x^7, c u, c u x, c u x^2
            uint32_t c64 = c;
#pragma unroll
            for (int b = 0; b < 6; b++) c64 = xtime16(c64);
            const uint32_t cu = times_u(c), cu2 = xtime16(cu);
            x[0][0] = x[1][0] = c64;
            x[0][1] = x[1][1] = xtime16(c64);
            x[0][2] = x[1][2] = cu;
            x[0][3] = x[1][3] = cu2;
            f[0] = 0u;
            f[1] = xtime16(cu2);
        } else if (wide) {  // bits 11..15: c u x^3 .. c u x^7
            uint32_t cu8 = times_u(c);
#pragma unroll
            for (int b = 0; b < 3; b++) cu8 = xtime16(cu8);
            run(cu8, x[0]);
            f[0] = 0u;
            f[1] = xtime16(x[0][3]);
            run(cu8, x[1]);
        } else {  // nibble fields 2q and 2q + 1
            const uint32_t v = q == 0u ? c : times_u(c);
            run(v, x[0]);
            run(xtime16(x[0][3]), x[1]);
            f[0] = f[1] = 0u;
        }
        uint32_t e[16];  // entries 2i, packed with entry 2i + 1: four 16-byte chunks
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                uint32_t v = f[h];
#pragma unroll
                for (int b = 1; b < 4; b++)
                    if ((i >> (b - 1)) & 1) v ^= x[h][b];
                e[8 * h + i] = v | ((v ^ x[h][0]) << 16);
            }
        // Neighbouring lanes build neighbouring tables, 256 or 128 bytes
        // apart: the same banks.  Store i of a lane writes chunk i ^ (lane
        // mod 4), so that one store spreads over four chunk positions
        // (16-way bank conflicts instead of 64-way).
        uint4 *d = reinterpret_cast<uint4 *>(dst);
        const uint32_t sw = tid & 3u;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t c = i ^ sw;
            // Chunk c by masks rather than selects, which the compiler turns
            // into an indexed read of e[] from scratch memory.
            const uint32_t m1 = 0u - (c & 1u), m2 = 0u - ((c >> 1) & 1u);
            uint32_t w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t lo = e[u] ^ ((e[u] ^ e[4 + u]) & m1), hi = e[8 + u] ^ ((e[8 + u] ^ e[12 + u]) & m1);
                w[u] = lo ^ ((lo ^ hi) & m2);
            }
            d[c] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    };
    auto build_tables = [&]() {
        if (inv || fwd) {
#pragma unroll
            for (uint32_t k = 0; k < 2; k++) {
                const uint32_t item = tid + k * nthr;
                if (item_node(item) == ~0u) continue;
                const uint32_t q = item >> lp, rest = item & (per - 1u), n = rest & (G - 1u);
                const uint32_t beta = g - 1u - (31u - static_cast<uint32_t>(__builtin_clz(G - 1u - n)));
                const uint32_t first = G - (G >> beta);
                build_item(fac[k], q, tab_n(beta) == kTabW,
                           ((!inv || rest >= G) ? fwd_t : inv_t) + level_off(beta) + (n - first) * tab_n(beta) + 32u * q);
            }
        }
        if (row_item) {
            const uint32_t q = tid >> g, off = (tid & (G - 1u)) * kTab + 32u * q;
            if (a.scale_in) build_item(fac_r[0], q, false, sin_t + off);
            if (a.scale_out) build_item(fac_r[1], q, false, sout_t + off);
            if (a.scale_der) build_item(fac_r[2], q, false, sder_t + off);
        }
    };
    auto stage_tables = [&]() {
        build_tables();
        if (der && !unit) stage_chunks(der_t, 16u * (kTab / 8u), [&](uint32_t c) {
            return reinterpret_cast<const uint4 *>(a.deriv_tabs) + c;
        });
        __syncthreads();
    };

    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63u;
    const uint64_t L = a.L;
    uint64_t cb = 0;
    bool valid = false;
    const uint64_t z = blockIdx.z;
    const uint32_t nst = (g + RB - 1) / RB;
    uint32_t v[R], nx[R];

    auto in_ptr = [&](uint32_t r) -> const uint8_t * {
        if (a.in_rows) return load_uniform(a.in_rows + r);
        return r < a.in_limit ? a.in_base + z * a.in_cw_stride + static_cast<uint64_t>(r) * L : nullptr;
    };
    auto out_ptr = [&](uint32_t r) -> uint8_t * {
        if (a.out_rows) return load_uniform(a.out_rows + r);
        return (r >= a.out_first && r < a.out_limit)
                   ? a.out_base + z * a.out_cw_stride + static_cast<uint64_t>(r - a.out_first) * L
                   : nullptr;
    };
    // The raw rows of column tile `tile` in the layout of stage (s_lo, t).
    auto load_raw = [&](uint32_t s_lo, uint32_t t, uint32_t tile, uint32_t(&dst)[R]) {
        const uint64_t cbx = (static_cast<uint64_t>(tile) * kTileDwords + lane) * 4u;
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint32_t p = rho<R>(wv, s_lo, t, j), r = base | (p << b_lo);
            const uint8_t *src = in_ptr(r);
            if constexpr (A4)
                dst[j] = ld_col4(src, a.node_fac, static_cast<uint32_t>(cbx), L);
            else
                dst[j] = (src && cbx < L) ? ld_col(src, cbx, L) : 0u;
        }
    };
    // Copies and input scaling of the loaded tile.
    auto prep_loaded = [&](uint32_t s_lo, uint32_t t) {
        if (!a.copy_rows && !a.scale_in) return;
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint32_t p = rho<R>(wv, s_lo, t, j), r = base | (p << b_lo);
            if (a.copy_rows) {
                uint8_t *cp = load_uniform(a.copy_rows + r);
                if constexpr (A4)
                    st_col4(cp, a.node_fac, static_cast<uint32_t>(cb), L, v[j]);
                else if (cp && valid)
                    st_col(cp, cb, L, v[j]);
            }
            if (a.scale_in) v[j] = gmul2(sin_t + p * kTab, v[j]);
        }
    };
    auto store_global = [&](uint32_t s_lo, uint32_t t) {
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint32_t p = rho<R>(wv, s_lo, t, j), r = base | (p << b_lo);
            uint8_t *dst = half && r >= a.row_limit ? nullptr : out_ptr(r);
            if constexpr (A4)
                st_col4(dst, a.node_fac, static_cast<uint32_t>(cb), L,
                        a.scale_out ? gmul2(sout_t + p * kTab, v[j]) : v[j]);
            else if (dst && valid)
                st_col(dst, cb, L, a.scale_out ? gmul2(sout_t + p * kTab, v[j]) : v[j]);
        }
    };
    auto load_lds = [&](uint32_t s_lo, uint32_t t) {
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = data[rho<R>(wv, s_lo, t, j) * kTileDwords + lane];
    };
    auto store_lds = [&](uint32_t s_lo, uint32_t t) {
#pragma unroll
        for (int j = 0; j < R; j++) data[rho<R>(wv, s_lo, t, j) * kTileDwords + lane] = v[j];
    };
    auto node_tab = [&](const uint16_t *tabs, uint32_t beta, uint32_t low_row) {
        return tabs + level_off(beta) + (low_row >> (beta + 1)) * tab_n(beta);
    };
    // A product by the node table of level beta, in that level's format.
    auto node_mul = [&](const uint16_t *tab, uint32_t beta, uint32_t x) {
        if (MX && beta == 0) return gmul2(tab, x);
        return gmul2n<W>(tab, x);
    };
    // In a stage over T bits at s_lo, T known at compile time, register j
    // holds local row (u mod 2^s_lo) | ((j mod 2^T) << s_lo) | ((u >> s_lo) << (s_lo + T))
    // with u = wv * (R >> T) + (j >> T).  The node of level s_lo + bs is the
    // level's first node plus ((u >> s_lo) << (T - 1 - bs)) plus the
    // compile-time ((j mod 2^T) >> (bs + 1)): one wave-uniform base per level
    // and value of j >> T, the rest in the LDS instruction's offset.
    auto level_base = [&](const uint16_t *tabs, uint32_t s_lo, int T, int bs, int uidx) -> uint32_t {
        const uint32_t beta = s_lo + static_cast<uint32_t>(bs);
        const uint32_t u = wv * static_cast<uint32_t>(R >> T) + static_cast<uint32_t>(uidx);
        const uint32_t node = (u >> s_lo) << (T - 1 - bs);  // within the level
        uint32_t lb = __builtin_amdgcn_readfirstlane(lds_byte(tabs) + (level_off(beta) + node * tab_n(beta)) * 2);
        asm("" : "+s"(lb));
        return lb;
    };
    // The first node of every level has the factor W(0) = 0.  In the pass's
    // top stage no pass bit lies above the stage, so a butterfly sits on its
    // level's first node exactly when the point base of the group has no bit
    // above the level and the register index none above the butterfly bit.
    // When the first holds for every level of the stage (uniform over the
    // workgroup), the stage runs in a form whose butterflies on the first
    // node skip their product (the reads and the combine) at compile time;
    // the XOR of the pair remains.
    auto zero_stage = [&](uint32_t pbase, uint32_t s_lo, uint32_t t) {
        return s_lo + t == g && (pbase >> (b_lo + s_lo + 1u)) == 0u;
    };
    // The butterflies of a level are software pipelined: the table reads of
    // product n are issued before product n - kPipe is combined.
    // fc: the stage starts at the pass's first level (s_lo == 0), whose
    // tables are nibble tables in the mixed format.
    // hc (forward, top stage): only the rows with the stage's top bit clear
    // are needed; the top level's high outputs and every butterfly of the
    // other half below it are skipped.
    // tp (kAbs kernels): the pass's top stage, whose node of register j at
    // level s_lo + bs is the level's table (j mod 2^T) >> (bs + 1) for every
    // wave: the table addresses are compile-time constants.
    auto stage_fast = [&](const uint16_t *tabs, uint32_t s_lo, auto tc, auto ic, auto zc, auto fc, auto hc, auto tp) {
        constexpr int T = decltype(tc)::value;
        constexpr bool inverse = decltype(ic)::value;
        constexpr bool HALF = decltype(hc)::value;
        constexpr bool ABS = kAbs && decltype(tp)::value;
        constexpr int NU = R >> T;
        static_for<T>([&](auto stepc) {
            constexpr int step = decltype(stepc)::value;
            constexpr int bs = inverse ? step : T - 1 - step;
            constexpr bool LW = W && !(MX && decltype(fc)::value && bs == 0);  // this level's format
            constexpr uint32_t kLN = LW ? kTabW : kTab;
            static_assert(!ABS || LW, "compile-time table addresses use the wide format");
            // The level's first table's LDS byte address (ABS).
            constexpr uint32_t kLevelByte =
                ABS ? (inverse ? 0u : kFwdByte) + level_off_bytes(kGc, GL - T + bs, MX, kTN) : 0u;
            uint32_t lb[NU];
            if constexpr (!ABS) {
#pragma unroll
                for (int q = 0; q < NU; q++) lb[q] = level_base(tabs, s_lo, T, bs, q);
            }
            // Wide levels below the top stage: address registers (above).
            constexpr bool BASE = !ABS && LW && kBaseRegs;
            WideAddr ad[BASE ? NU : 1];
            if constexpr (BASE) {
#pragma unroll
                for (int q = 0; q < NU; q++) wide_addr_init(ad[q], lb[q]);
            }
            // Butterfly n of the level: registers jn and jn + 2^bs.
            auto jn = [&](int n) { return ((n >> bs) << (bs + 1)) | (n & ((1 << bs) - 1)); };
            auto off = [&](int j) { return static_cast<uint32_t>((j & ((1 << T) - 1)) >> (bs + 1)) * (kLN * 2); };
            auto zero = [&](int j) { return decltype(zc)::value && off(j) == 0; };
            auto dead = [&](int j) { return HALF && bs < T - 1 && (j & (1 << (T - 1))) != 0; };
            Parts<LW> pp[kPipe + 1];
#pragma unroll
            for (int n = 0; n < R / 2 + kPipe; n++) {
                if (n < R / 2) {
                    const int j = jn(n);
                    if constexpr (inverse) v[j + (1 << bs)] ^= v[j];
                    if (!zero(j) && !dead(j)) {
                        if constexpr (ABS)
                            gmul2_issue_abs(pp[n % (kPipe + 1)], kLevelByte + off(j), v[j + (1 << bs)]);
                        else if constexpr (BASE)
                            gmul2_issue_base(pp[n % (kPipe + 1)], ad[j >> T], off(j), v[j + (1 << bs)]);
                        else
                            gmul2_issue<LW>(pp[n % (kPipe + 1)], lb[j >> T], off(j), v[j + (1 << bs)]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                if (n >= kPipe) {
                    const int j = jn(n - kPipe);
                    if (!zero(j) && !dead(j)) v[j] ^= gmul2_combine<LW>(pp[(n - kPipe) % (kPipe + 1)]);
                    if constexpr (!inverse)
                        if (!dead(j) && !(HALF && bs == T - 1)) v[j + (1 << bs)] ^= v[j];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        });
    };
    // Full stages, and 3-bit stages (a 7-level pass), take the fast path.
    // The 8-level narrow-table form keeps one form of its stages: with both,
    // it needs more than the 128 registers of 4 waves per SIMD.
    auto fast_stage = [&](const uint16_t *tabs, uint32_t s_lo, uint32_t t, bool zs, bool hs, auto ic) {
        auto run = [&](auto tc) {
            auto go = [&](auto zc, auto fc) {
                // The top stage of a kAbs kernel: compile-time table addresses.
                auto go2 = [&](auto hc) {
                    if constexpr (kAbs && !decltype(fc)::value) {
                        if (s_lo + decltype(tc)::value == GL) {
                            stage_fast(tabs, s_lo, tc, ic, zc, fc, hc, std::true_type{});
                            return;
                        }
                    }
                    stage_fast(tabs, s_lo, tc, ic, zc, fc, hc, std::false_type{});
                };
                if constexpr (!decltype(ic)::value && (W || GL == 0)) {
                    if (hs) {
                        go2(std::true_type{});
                        return;
                    }
                }
                go2(std::false_type{});
            };
            if (MX && s_lo == 0) {
                if (zs)
                    go(std::true_type{}, std::bool_constant<MX>{});
                else
                    go(std::false_type{}, std::bool_constant<MX>{});
            } else if (zs && (W || GL == 0)) {
                go(std::true_type{}, std::false_type{});
            } else {
                go(std::false_type{}, std::false_type{});
            }
        };
        if (t == RB) {
            run(std::integral_constant<int, static_cast<int>(RB)>{});
            return true;
        }
        if constexpr ((GL == 0 || GL % 4 == 3) && RB == 4) {
            if (t == 3) {
                run(std::integral_constant<int, 3>{});
                return true;
            }
        }
        return false;
    };
    const uint32_t inv_pbase = a.inv_base + base, fwd_pbase = a.fwd_base + base;
    auto inv_stage = [&](uint32_t s_lo, uint32_t t) {
        if (fast_stage(inv_t, s_lo, t, zero_stage(inv_pbase, s_lo, t), false, std::true_type{})) return;
#pragma unroll
        for (int bs = 0; bs < static_cast<int>(RB); bs++) {
            if (static_cast<uint32_t>(bs) >= t) break;
#pragma unroll
            for (int j = 0; j < R; j++) {
                if (j & (1 << bs)) continue;
                const uint16_t *tab = node_tab(inv_t, s_lo + bs, rho<R>(wv, s_lo, t, j));
                v[j + (1 << bs)] ^= v[j];
                v[j] ^= node_mul(tab, s_lo + bs, v[j + (1 << bs)]);
            }
        }
    };
    auto fwd_stage = [&](uint32_t s_lo, uint32_t t) {
        // Half-output passes: below the top stage a wave's rows share the
        // pass's top bit when a stage tile is one wave's (R >> t == 1), and
        // the waves of the unneeded half have nothing to compute.
        if (half && s_lo + t < g && (R >> t) == 1 && ((wv >> (g - 1u - t)) & 1u)) return;
        if (fast_stage(fwd_t, s_lo, t, zero_stage(fwd_pbase, s_lo, t), half && s_lo + t == g, std::false_type{}))
            return;
#pragma unroll
        for (int bs = static_cast<int>(RB) - 1; bs >= 0; bs--) {
            if (static_cast<uint32_t>(bs) >= t) continue;
#pragma unroll
            for (int j = 0; j < R; j++) {
                if (j & (1 << bs)) continue;
                const uint16_t *tab = node_tab(fwd_t, s_lo + bs, rho<R>(wv, s_lo, t, j));
                v[j] ^= node_mul(tab, s_lo + bs, v[j + (1 << bs)]);
                v[j + (1 << bs)] ^= v[j];
            }
        }
    };
    // With the level count fixed, the first stage takes the remainder of g
    // over RB bits, so that the top stage is a full one: a 7-level pass runs
    // 3 levels, then 4 at compile-time table addresses (kAbs).  Otherwise the
    // last stage takes it (the register budget of the general kernels).
    constexpr bool kShortFirst = GL != 0 && GL % RB != 0;
    constexpr uint32_t kT0 = kShortFirst ? GL % RB : RB;
    auto stage_lo = [&](uint32_t k) { return kShortFirst ? (k == 0u ? 0u : kT0 + (k - 1u) * RB) : k * RB; };
    auto stage_t = [&](uint32_t k) { return kShortFirst ? (k == 0u ? kT0 : RB) : min(RB, g - k * RB); };

    const uint32_t tiles = static_cast<uint32_t>((L + kTileDwords * 4 - 1) / (kTileDwords * 4));
    const uint32_t t_first = blockIdx.x * tpb, t_end = min(tiles, t_first + tpb);
    const uint32_t k_in = inv ? 0u : nst - 1;  // stage layout the rows load in
    // The derivative starts in the last stage's layout and, with two stages,
    // ends in the first one's: either way the accumulator rows are added in
    // stage layout 0.
    const uint32_t k_der = 0;
    // The node factors load first, then the first tile's rows, and the
    // tables build and stage while the rows arrive (no row load depends on
    // a table).
    load_factors();
    if (t_first < t_end) load_raw(stage_lo(k_in), stage_t(k_in), t_first, nx);
    stage_tables();
    uint32_t ax[R];
    for (uint32_t tile = t_first; tile < t_end; tile++) {
        cb = (static_cast<uint64_t>(tile) * kTileDwords + lane) * 4u;
        valid = cb < L;
        // The derivative accumulator of this tile loads now; its latency
        // hides behind the stages that run before it is added.
        if (der && a.aux_base) {
#pragma unroll
            for (int j = 0; j < R; j++) {
                const uint32_t p = rho<R>(wv, stage_lo(k_der), stage_t(k_der), j);
                const uint8_t *row = a.aux_base + z * a.aux_cw_stride + static_cast<uint64_t>(base | (p << b_lo)) * L;
                if constexpr (A4)
                    ax[j] = ld_col4(row, a.node_fac, static_cast<uint32_t>(cb), L);
                else
                    ax[j] = valid ? ld_col(row, cb, L) : 0u;
            }
        }
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = nx[j];
        if (tile + 1 < t_end) load_raw(stage_lo(k_in), stage_t(k_in), tile + 1, nx);
        prep_loaded(stage_lo(k_in), stage_t(k_in));
        if (tile != t_first) __syncthreads();  // the previous tile is done with the LDS tile
        int cur = static_cast<int>(k_in);
        if (inv) {
            for (uint32_t k = 0; k < nst; k++) {
                if (k > 0) {
                    store_lds(stage_lo(cur), stage_t(cur));
                    __syncthreads();
                    cur = k;
                    load_lds(stage_lo(cur), stage_t(cur));
                }
                inv_stage(stage_lo(k), stage_t(k));
            }
        }
        if (der) {
            // D[p] = sum over pass bits beta clear in p of d_(b_lo+beta) * C[p | 2^beta].
            // In a stage layout the stage's bits of a register's row are
            // bits of the register index, so each stage's terms pair
            // registers j and j | 2^bs with no branch and no LDS read.  The
            // current stage is summed in place; each other stage is summed
            // in its own layout and brought back through the LDS tile.
            // Source-major: each coefficient c[i] is scaled by the table of
            // every stage bit set in i and added to the row with that bit
            // clear, so its nibble offsets serve all of its products (the
            // tables of one stage are consecutive).  With unit factors the
            // terms are plain XORs.
            auto der_stage = [&](const uint32_t(&c)[R], uint32_t(&d)[R], uint32_t s_lo, uint32_t t) {
#pragma unroll
                for (int j = 0; j < R; j++) d[j] = 0;
                if (unit) {
#pragma unroll
                    for (int i = 1; i < R; i++)
#pragma unroll
                        for (int bs = 0; bs < static_cast<int>(RB); bs++)
                            if (((i >> bs) & 1) && static_cast<uint32_t>(bs) < t) d[i ^ (1 << bs)] ^= c[i];
                    return;
                }
                const uint16_t *tab0 = der_t + (b_lo + s_lo) * kTab;
#pragma unroll
                for (int i = 1; i < R; i++) {
#pragma unroll
                    for (int bs = 0; bs < static_cast<int>(RB); bs++)
                        if ((i >> bs) & 1)
                            if (static_cast<uint32_t>(bs) < t) d[i ^ (1 << bs)] ^= gmul2(tab0 + bs * kTab, c[i]);
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            uint32_t acc[R];
            der_stage(v, acc, stage_lo(cur), stage_t(cur));
            __builtin_amdgcn_sched_barrier(0);
            // A pass has at most two stages (g <= 8 bits, 4 bits per stage
            // at R = 16, one stage below).  With two, the other stage's
            // terms are summed in its own layout: the coefficients go
            // through the tile into v in that layout, this stage's partial
            // sums take their place in the tile, and the result continues
            // in the other layout.
            if (nst == 2) {
                const uint32_t k = 1u - static_cast<uint32_t>(cur);
                store_lds(stage_lo(cur), stage_t(cur));
                __syncthreads();
                load_lds(stage_lo(k), stage_t(k));
                __syncthreads();
#pragma unroll
                for (int j = 0; j < R; j++) data[rho<R>(wv, stage_lo(cur), stage_t(cur), j) * kTileDwords + lane] = acc[j];
                der_stage(v, acc, stage_lo(k), stage_t(k));
                __builtin_amdgcn_sched_barrier(0);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < R; j++) acc[j] ^= data[rho<R>(wv, stage_lo(k), stage_t(k), j) * kTileDwords + lane];
                __syncthreads();  // the tile is free for the next layout change
                cur = static_cast<int>(k);
            }
            // Here cur == k_der: the accumulator rows loaded in this layout.
#pragma unroll
            for (int j = 0; j < R; j++) {
                const uint32_t x = a.aux_base ? acc[j] ^ ax[j] : acc[j];
                v[j] = a.scale_der ? gmul2(sder_t + rho<R>(wv, stage_lo(cur), stage_t(cur), j) * kTab, x) : x;
            }
        }
        if (fwd) {
            for (int k = static_cast<int>(nst) - 1; k >= 0; k--) {
                if (cur != k) {
                    store_lds(stage_lo(cur), stage_t(cur));
                    __syncthreads();
                    cur = k;
                    load_lds(stage_lo(cur), stage_t(cur));
                }
                fwd_stage(stage_lo(k), stage_t(k));
            }
        }
        store_global(stage_lo(cur), stage_t(cur));
    }
}

template <int R, bool A4, uint32_t GL = 0, bool W = false, uint32_t FX = kFxDynamic, bool MX = false>
hipError_t launch_pass_r(const AfftPassArgs &a, uint32_t D, uint32_t count, size_t lds, hipStream_t s) {
    const uint32_t G = 1u << a.g;
    const uint64_t tiles = (a.L + kTileDwords * 4 - 1) / (kTileDwords * 4);
    // Groups whose rows all lie at or above row_limit are not launched:
    // group gidx starts at row ((gidx >> b_lo) << (b_lo + g)) | (gidx & (2^b_lo - 1)),
    // so the live groups are the first ((row_limit - 1) >> (b_lo + g)) + 1 blocks of 2^b_lo.
    uint64_t groups = uint64_t{1} << (D - a.g);
    if (a.row_limit)
        groups = std::min<uint64_t>(groups, ((uint64_t{(a.row_limit - 1u) >> (a.b_lo + a.g)}) + 1) << a.b_lo);
    // Tiles per workgroup: enough workgroups to fill every CU to its LDS
    // capacity, no more, so each group's tables are built as few times as
    // possible.
    const uint64_t per_cu = lds ? std::max<uint64_t>(1, (160u * 1024u) / lds) : 8;
    const uint64_t want = tiles * groups * count / (256 * per_cu);
    const uint32_t tpb = static_cast<uint32_t>(want < 1 ? 1 : (want > tiles ? tiles : want));
    const dim3 grid(static_cast<uint32_t>((tiles + tpb - 1) / tpb), static_cast<uint32_t>(groups), count);
    const dim3 block((G / R) * 64);
    const void *fn = reinterpret_cast<const void *>(&afft_pass_kernel<R, A4, GL, W, MX, FX>);
    hipError_t err = lds_opt_in(fn, 160 * 1024);
    if (err != hipSuccess) return err;
    if constexpr (W && GL != 0 && FX != kFxDynamic) {
        // Constant top-stage table addresses assume the dynamic LDS block
        // starts at address 0: the kernel declares no static LDS.
        static const hipError_t placed = [fn] {
            hipFuncAttributes attr{};
            const hipError_t e = hipFuncGetAttributes(&attr, fn);
            return e != hipSuccess ? e : (attr.sharedSizeBytes == 0 ? hipSuccess : hipErrorInvalidDeviceFunction);
        }();
        if (placed != hipSuccess) return placed;
    }
    hipLaunchKernelGGL((afft_pass_kernel<R, A4, GL, W, MX, FX>), grid, block, lds, s, a, tpb);
    return hipGetLastError();
}

// ---- constant tables -------------------------------------------------------


__global__ __launch_bounds__(256) void afft_tables_kernel(const uint16_t *c, uint32_t count, uint16_t *out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint32_t basis[16];
    basis[0] = c[i];
    for (int b = 1; b < 8; b++) basis[b] = xtime16(basis[b - 1]);
    basis[8] = times_u(c[i]);
    for (int b = 9; b < 16; b++) basis[b] = xtime16(basis[b - 1]);
    uint32_t *o = reinterpret_cast<uint32_t *>(out + static_cast<uint64_t>(i) * kTab);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int n = 0; n < 16; n += 2) {
            uint32_t e0 = 0, e1 = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                if ((n >> b) & 1) e0 ^= basis[4 * q + b];
                if (((n + 1) >> b) & 1) e1 ^= basis[4 * q + b];
            }
            o[(q * 16 + n) / 2] = e0 | (e1 << 16);
        }
}

// Wide tables (the 6 + 5 + 5 split of gmulw).
// ---- locator (afft/decoder.rs:672-771) -------------------------------------

__device__ __forceinline__ uint32_t addmod(uint32_t x, uint32_t y) {
    const uint32_t s = x + y;
    return s >= kAfftOrder ? s - kAfftOrder : s;
}
__device__ __forceinline__ uint32_t submod(uint32_t x, uint32_t y) { return x >= y ? x - y : x + kAfftOrder - y; }
__device__ __forceinline__ uint32_t mulmod(uint32_t x, uint32_t y) {
    const uint32_t p = x * y;  // x, y < 65535: p < 2^32
    uint32_t t = (p & 0xffffu) + (p >> 16);
    t = (t & 0xffffu) + (t >> 16);
    return t >= kAfftOrder ? t - kAfftOrder : t;
}

enum LocMode : uint32_t {
    kLocInit = 1,      // start from the unknown-point indicator
    kLocMulWht = 2,    // multiply by the WHT of log and transform again
    kLocFinal = 4,     // scale by 2^-D and exponentiate
    kLocHiLayout = 8,  // 16 columns x 2^H rows of stride 1024
};

// Block-local Walsh-Hadamard transform mod 65535 over LDS index bits
// [lo, lo + bits).
__device__ void lds_wht(uint32_t *s, uint32_t lo, uint32_t bits) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = lo; b < lo + bits; b++) {
        __syncthreads();
        const uint32_t h = 1u << b;
        if (!(tid & h) && (tid | h) < blockDim.x) {
            const uint32_t x = s[tid], y = s[tid | h];
            s[tid] = addmod(x, y);
            s[tid | h] = submod(x, y);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(1024) void afft_locator_kernel(AfftLocatorArgs a, uint32_t mode, uint32_t wht_lo,
                                                            uint32_t wht_bits, uint32_t inv_n) {
    __shared__ uint32_t s[1024];
    const uint32_t tid = threadIdx.x;
    uint32_t x;
    if (mode & kLocHiLayout) x = blockIdx.x * 16u + (tid & 15u) + ((tid >> 4) << 10);
    else x = blockIdx.x * blockDim.x + tid;
    const bool known = (a.known_bits[x >> 5] >> (x & 31u)) & 1u;
    s[tid] = (mode & kLocInit) ? (known ? 0u : 1u) : a.work[x];
    lds_wht(s, wht_lo, wht_bits);
    if (mode & kLocMulWht) {
        s[tid] = mulmod(s[tid], a.log_wht[x]);
        lds_wht(s, wht_lo, wht_bits);
    }
    if (mode & kLocFinal) {
        const uint32_t e = mulmod(s[tid], inv_n);
        a.out[x] = known ? a.exp[e] : a.exp[e ? kAfftOrder - e : 0];
    } else {
        a.work[x] = s[tid];
    }
}

__global__ __launch_bounds__(256) void afft_rows_kernel(AfftRowsArgs a) {
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= (1u << a.D)) return;
    const bool known = x < a.n && ((a.known_bits[x >> 5] >> (x & 31u)) & 1u);
    const uint8_t *src = nullptr;
    if (known && a.data_base) src = x < a.k ? a.data_base + x * a.L : a.repair_base + (x - a.k) * a.L;
    a.in_rows[x] = src;
    a.copy_rows[x] = (known && x < a.k) ? a.out + x * a.L : nullptr;
    a.out_rows[x] = (!known && x < a.k) ? a.out + x * a.L : nullptr;
}

__global__ __launch_bounds__(256) void afft_scatter_kernel(const uint16_t *idx, const uint8_t *const *ptrs,
                                                           uint32_t count, const uint8_t **in_rows) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < count) in_rows[idx[t]] = ptrs[t];
}

}  // namespace

size_t afft_pass_lds_bytes(uint32_t g, uint32_t flags, bool scale_in, bool scale_out, bool scale_der, bool wide,
                           bool mixed) {
    const size_t G = size_t{1} << g, tab = kTab * 2, ntab = (wide ? kTabW : kTab) * 2;
    // Mixed: the first level's G/2 tables in the nibble format.
    const size_t nodes = mixed && wide ? (G / 2) * tab + (G / 2 - 1) * ntab : (G - 1) * ntab;
    size_t b = G * kTileDwords * 4;
    if (flags & kAfftInv) b += nodes;
    if (flags & kAfftFwd) b += nodes;
    if ((flags & kAfftDeriv) && !(flags & kAfftDerivUnit)) b += 16 * tab;
    if (scale_in) b += G * tab;
    if (scale_out) b += G * tab;
    if (scale_der) b += G * tab;
    return b;
}

// The kFx* feature set of a pass's arguments.
uint32_t pass_features(const AfftPassArgs &a) {
    return ((a.flags & kAfftInv) ? kFxInv : 0u) | ((a.flags & kAfftFwd) ? kFxFwd : 0u) |
           ((a.flags & kAfftDeriv) ? kFxDer : 0u) | ((a.flags & kAfftDerivUnit) ? kFxUnit : 0u) |
           (a.scale_in ? kFxSin : 0u) | (a.scale_out ? kFxSout : 0u) | (a.scale_der ? kFxSder : 0u) |
           (a.copy_rows ? kFxCopy : 0u) | (a.aux_base ? kFxAux : 0u) | (a.in_rows ? kFxInRows : 0u) |
           (a.out_rows ? kFxOutRows : 0u);
}

hipError_t launch_afft_pass(const AfftPassArgs &a, uint32_t D, uint32_t count, hipStream_t s) {
    if (a.g == 0 || a.g > 8 || a.b_lo + a.g > D || D > kAfftMaxLog || count == 0 || count > 65535 || a.L == 0)
        return hipErrorInvalidValue;
    if (a.scale_der && !(a.flags & (kAfftDeriv | kAfftDerivUnit))) return hipErrorInvalidValue;
    if ((a.flags & (kAfftInv | kAfftFwd)) && !a.node_fac) return hipErrorInvalidValue;
    const bool sin = a.scale_in != nullptr, sout = a.scale_out != nullptr, sder = a.scale_der != nullptr;
    const size_t lds = afft_pass_lds_bytes(a.g, a.flags, sin, sout, sder, false);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const size_t lds_w = afft_pass_lds_bytes(a.g, a.flags, sin, sout, sder, true);
    const size_t lds_m = afft_pass_lds_bytes(a.g, a.flags, sin, sout, sder, true, true);
    // Dword buffer instructions when every row the pass touches starts on a
    // multiple of 4 (row tables: as declared by the caller).
    auto al4 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; };
    const bool tables = a.in_rows || a.copy_rows || a.out_rows;
    const bool a4 = a.L % 4 == 0 && a.L <= 0x7fffffffull && al4(a.in_base) && al4(a.out_base) && al4(a.aux_base) &&
                    (a.in_cw_stride | a.out_cw_stride | a.aux_cw_stride) % 4 == 0 && (!tables || a.rows_aligned4);
    // Wide node tables for the 16-row passes whose tables still fit, unless
    // the nibble tables' smaller footprint lets more waves share a CU (at
    // most 16 at these kernels' register use: 4 per SIMD).
    auto waves_per_cu = [&](size_t bytes) {
        const size_t wgs = (160u * 1024u) / bytes, waves = (size_t{1} << a.g) / 16;
        return std::min<size_t>(16, wgs * waves);
    };
    const uint32_t fx = pass_features(a);
    if (a4 && a.g >= 4 && lds_w <= 160 * 1024 && waves_per_cu(lds_w) >= waves_per_cu(lds)) {
        if (a.g == 8) {
            // The feature sets of the encode and decode pass chains.
            switch (fx) {
                case kFxInv: return launch_pass_r<16, true, 8, true, kFxInv>(a, D, count, lds_w, s);
                case kFxFwd: return launch_pass_r<16, true, 8, true, kFxFwd>(a, D, count, lds_w, s);
                case kFxInv | kFxInRows | kFxCopy | kFxSin:
                    return launch_pass_r<16, true, 8, true, kFxInv | kFxInRows | kFxCopy | kFxSinRt>(a, D, count, lds_w,
                                                                                                   s);
                case kFxInv | kFxSout: return launch_pass_r<16, true, 8, true, kFxInv | kFxSout>(a, D, count, lds_w, s);
                case kFxUnit: return launch_pass_r<16, true, 8, true, kFxUnit>(a, D, count, lds_w, s);
                case kFxUnit | kFxFwd | kFxSder | kFxAux:
                    return launch_pass_r<16, true, 8, true, kFxUnit | kFxFwd | kFxSder | kFxAux>(a, D, count, lds_w, s);
                case kFxFwd | kFxOutRows | kFxSout:
                    return launch_pass_r<16, true, 8, true, kFxFwd | kFxOutRows | kFxSout>(a, D, count, lds_w, s);
                default: return launch_pass_r<16, true, 8, true>(a, D, count, lds_w, s);
            }
        }
        return launch_pass_r<16, true, 0, true>(a, D, count, lds_w, s);
    }
    // Mixed tables where they keep the nibble tables' waves per CU: three
    // reads per element above the pass's first level instead of four.
    if (a4 && a.g >= 4 && a.g < 8 && lds_m <= 160 * 1024 && waves_per_cu(lds_m) >= waves_per_cu(lds)) {
        if (fx == (kFxInv | kFxFwd) && a.g == 7)
            return launch_pass_r<16, true, 7, true, kFxInv | kFxFwd, true>(a, D, count, lds_m, s);
        return launch_pass_r<16, true, 0, true, kFxDynamic, true>(a, D, count, lds_m, s);
    }
    if (a4) {
        switch (a.g) {
            case 8: return launch_pass_r<16, true, 8>(a, D, count, lds, s);
            case 1: return launch_pass_r<2, true>(a, D, count, lds, s);
            case 2: return launch_pass_r<4, true>(a, D, count, lds, s);
            case 3: return launch_pass_r<8, true>(a, D, count, lds, s);
            default:
                if (fx == (kFxInv | kFxFwd)) return launch_pass_r<16, true, 0, false, kFxInv | kFxFwd>(a, D, count, lds, s);
                return launch_pass_r<16, true>(a, D, count, lds, s);
        }
    }
    switch (a.g) {
        case 1: return launch_pass_r<2, false>(a, D, count, lds, s);
        case 2: return launch_pass_r<4, false>(a, D, count, lds, s);
        case 3: return launch_pass_r<8, false>(a, D, count, lds, s);
        default: return launch_pass_r<16, false>(a, D, count, lds, s);
    }
}

hipError_t launch_afft_tables(const uint16_t *c, uint32_t count, uint16_t *out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(afft_tables_kernel, dim3((count + 255) / 256), dim3(256), 0, s, c, count, out);
    return hipGetLastError();
}

hipError_t launch_afft_locator(const AfftLocatorArgs &a, hipStream_t s) {
    const uint32_t D = a.D;
    if (D == 0 || D > kAfftMaxLog) return hipErrorInvalidValue;
    uint32_t inv_n = 1;
    for (uint32_t i = 0; i < D; i++) inv_n = static_cast<uint32_t>((static_cast<uint64_t>(inv_n) * 32768u) % kAfftOrder);
    if (D <= 10) {
        hipLaunchKernelGGL(afft_locator_kernel, dim3(1), dim3(1u << D), 0, s, a, kLocInit | kLocMulWht | kLocFinal,
                           0u, D, inv_n);
        return hipGetLastError();
    }
    const uint32_t H = D - 10;
    hipLaunchKernelGGL(afft_locator_kernel, dim3(1u << H), dim3(1024), 0, s, a, kLocInit, 0u, 10u, inv_n);
    hipLaunchKernelGGL(afft_locator_kernel, dim3(1024 / 16), dim3(16u << H), 0, s, a, kLocMulWht | kLocHiLayout, 4u,
                       H, inv_n);
    hipLaunchKernelGGL(afft_locator_kernel, dim3(1u << H), dim3(1024), 0, s, a, kLocFinal, 0u, 10u, inv_n);
    return hipGetLastError();
}

hipError_t launch_afft_rows(const AfftRowsArgs &a, hipStream_t s) {
    const uint32_t n = 1u << a.D;
    hipLaunchKernelGGL(afft_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_afft_scatter_ptrs(const uint16_t *indices, const uint8_t *const *ptrs, uint32_t count,
                                    const uint8_t **in_rows, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(afft_scatter_kernel, dim3((count + 255) / 256), dim3(256), 0, s, indices, ptrs, count,
                       in_rows);
    return hipGetLastError();
}

}  // namespace scrs
