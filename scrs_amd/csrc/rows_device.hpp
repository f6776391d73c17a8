// rows_device.hpp - device building blocks of the table multiply-XOR kernels:
// 16-byte column loads and stores, LDS product-table entries and the
// per-term fold of the GF(2^8) and GF(2^16) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scrs {
namespace rowsdev {

// Payload rows always live in global memory (device or mapped host
// allocations).  Pointers read from tables are generic to the compiler,
// which would issue flat instructions for them; a flat access counts
// against the LDS wait counter too, so a wait for a table lookup would also
// wait for every source load still in flight.  Accesses go through
// global-address-space pointers instead.
#define SCRS_GLOBAL __attribute__((address_space(1)))
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ const SCRS_GLOBAL T *gbl(const void *p) {
    return (const SCRS_GLOBAL T *)(p);
}
template <typename T>
__device__ __forceinline__ SCRS_GLOBAL T *gbl_mut(void *p) {
    return (SCRS_GLOBAL T *)(p);
}

// 16 bytes at column p of a row (p < L, p a multiple of 16).  VEC: the row
// address and L are multiples of 16, known at compile time.  Otherwise the
// width follows the row's alignment at run time (row pointers are
// wave-uniform, so the branch is too): one 16-byte load when the address and
// L are multiples of 16; else the five aligned dwords covering the 16 bytes,
// funnel-shifted into place.  An aligned dword holding at least one byte of
// the row lies in the same page as that byte, so the loads never leave
// mapped memory; bytes past L are don't-cares (never stored).
template <bool VEC>
__device__ __forceinline__ void load16(const uint8_t *src, uint64_t p, uint64_t L, uint32_t (&w)[4]) {
    if (VEC || (((reinterpret_cast<uint64_t>(src) | L) & 15) == 0)) {
        const v4u32 v = *gbl<v4u32>(src + p);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {
        const uint8_t *at = src + p;
        const uint32_t sh = static_cast<uint32_t>(reinterpret_cast<uint64_t>(at) & 3u);
        const uint8_t *al = at - sh;
        const uint8_t *end = src + L;
        uint32_t d[5];
#pragma unroll
        for (int i = 0; i < 5; i++) d[i] = al + 4 * i < end ? *gbl<uint32_t>(al + 4 * i) : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = __builtin_amdgcn_alignbyte(d[j + 1], d[j], sh);
    }
}

// Byte stores of 16 bytes at column p, none at or past L.  Out of line: the
// unaligned-row case is rare and its code would otherwise inflate the
// register budget of every kernel that inlines store16.
__device__ __noinline__ void store16_bytes(uint8_t *dst, uint64_t p, uint64_t L, uint32_t w0, uint32_t w1,
                                            uint32_t w2, uint32_t w3) {
    const uint32_t w[4] = {w0, w1, w2, w3};
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint64_t q = p + j * 4 + b;
            if (q < L) *gbl_mut<uint8_t>(dst + q) = static_cast<uint8_t>(w[j] >> (8 * b));
        }
}

// Store counterpart of load16: 16-byte or dword stores where the row's
// alignment allows, bytes otherwise; nothing at or past L is written.
template <bool VEC>
__device__ __forceinline__ void store16(uint8_t *dst, uint64_t p, uint64_t L, const uint32_t (&w)[4]) {
    const uint64_t al = reinterpret_cast<uint64_t>(dst) | L;
    if (VEC || (al & 15) == 0) {
        *gbl_mut<v4u32>(dst + p) = v4u32{w[0], w[1], w[2], w[3]};
    } else if ((al & 3) == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (p + 4 * j < L) *gbl_mut<uint32_t>(dst + p + 4 * j) = w[j];
    } else {
        store16_bytes(dst, p, L, w[0], w[1], w[2], w[3]);
    }
}

// XOR 16 bytes into a row whose address and L are multiples of 4 (term-split
// partial sums): one atomic per dword inside the row.
__device__ __forceinline__ void xor16_atomic(uint8_t *dst, uint64_t p, uint64_t L, const uint32_t (&w)[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (p + 4 * j < L)
            __hip_atomic_fetch_xor(gbl_mut<uint32_t>(dst + p + 4 * j), w[j], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
}

// Entry of RW bytes at byte offset `off` of the LDS chunk.
template <int RW> struct Entry;
template <> struct Entry<16> {
    uint32_t d[4];
    __device__ __forceinline__ void load(const uint8_t *lds, uint32_t off) {
        const uint4 v = *reinterpret_cast<const uint4 *>(lds + off);
        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
};
template <> struct Entry<8> {
    uint32_t d[2];
    __device__ __forceinline__ void load(const uint8_t *lds, uint32_t off) {
        const uint2 v = *reinterpret_cast<const uint2 *>(lds + off);
        d[0] = v.x; d[1] = v.y;
    }
};
template <> struct Entry<4> {
    uint32_t d[1];
    __device__ __forceinline__ void load(const uint8_t *lds, uint32_t off) {
        d[0] = *reinterpret_cast<const uint32_t *>(lds + off);
    }
};

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Byte B of x, zero-extended, AND m: one SDWA instruction (src0_sel BYTE_B).
// Written out because the compiler folds only some shift-and-mask pairs into
// SDWA operands.  A VALU instruction, no memory access.
template <int B>
__device__ __forceinline__ uint32_t byte_and(uint32_t x, uint32_t m) {
    uint32_t r;
    if constexpr (B == 0) r = x & m;
    else if constexpr (B == 1)
        asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 src1_sel:DWORD"
            : "=v"(r) : "v"(x), "s"(m));
    else if constexpr (B == 2)
        asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2 src1_sel:DWORD"
            : "=v"(r) : "v"(x), "s"(m));
    else
        asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_3 src1_sel:DWORD"
            : "=v"(r) : "v"(x), "s"(m));
    return r;
}

// One source byte (byte B of the dword whose rotated copies are lo and hi)
// folded into its 16 row accumulators.
template <int RW, int B>
__device__ __forceinline__ void fold_byte(uint32_t (&acc)[RW / 4], const uint8_t *tab, uint32_t lo, uint32_t hi) {
    constexpr int SH = RW == 16 ? 4 : (RW == 8 ? 3 : 2);  // log2(RW)
    constexpr uint32_t BMASK = 0x0Fu << SH;                 // one entry offset
    Entry<RW> el, eh;
    if constexpr (SH == 4) {
        el.load(tab, byte_and<B>(lo, BMASK));
        eh.load(tab + 16 * RW, byte_and<B>(hi, BMASK));
    } else {
        el.load(tab, (lo >> (8 * B)) & BMASK);
        eh.load(tab + 16 * RW, (hi >> (8 * B)) & BMASK);
    }
#pragma unroll
    for (int d = 0; d < RW / 4; d++) acc[d] = xor3(acc[d], el.d[d], eh.d[d]);
}

// Fold one 16-byte source slice into the accumulators.  `tab` points at the
// term's 32 entries (lo 0..15, hi 16..31) of RW bytes each.
template <int RW>
__device__ __forceinline__ void fold_term(uint32_t (&acc)[16][RW / 4], const uint8_t *tab, const uint32_t (&w)[4]) {
    constexpr int SH = RW == 16 ? 4 : (RW == 8 ? 3 : 2);  // log2(RW)
#pragma unroll
    for (int j = 0; j < 4; j++) {
        // Rotations place each byte's low (high) nibble times RW into bits
        // SH..SH+3 of that byte; masking byte b gives the entry offset.  With
        // 16-byte entries (SH = 4) the high nibble is in place already and a
        // left shift by 4 places the low one, and each offset is one SDWA
        // AND of byte b (byte_and).
        const uint32_t lo = SH == 4 ? w[j] << 4 : __builtin_amdgcn_alignbit(w[j], w[j], 32 - SH);
        const uint32_t hi = SH == 4 ? w[j] : __builtin_amdgcn_alignbit(w[j], w[j], 4 - SH);
        fold_byte<RW, 0>(acc[4 * j + 0], tab, lo, hi);
        fold_byte<RW, 1>(acc[4 * j + 1], tab, lo, hi);
        fold_byte<RW, 2>(acc[4 * j + 2], tab, lo, hi);
        fold_byte<RW, 3>(acc[4 * j + 3], tab, lo, hi);
        // Fold four columns at a time: keeps the lookups in flight bounded
        // and the register budget at four waves per SIMD.
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Entries of the per-codeword pointer / row-count tables are the same for
// every lane and the kernel never writes those tables: read them through the
// constant address space so they become scalar loads into SGPRs.
typedef __attribute__((address_space(4))) const uint64_t *const_u64_ptr;
typedef __attribute__((address_space(4))) const uint32_t *const_u32_ptr;

template <typename T>
__device__ __forceinline__ T *load_uniform(T *const *ptr) {
    return reinterpret_cast<T *>(*(const_u64_ptr)(ptr));
}
__device__ __forceinline__ uint32_t load_uniform(const uint32_t *ptr) { return *(const_u32_ptr)(ptr); }

// ---- GF(2^16): entries of EB = 2 * rows bytes, 64 entries per term ----

// LDS dword index of dword d of entry e = 16q + n (nibble position q, value
// n) of one term's staged table.  32-byte entries are split into two 16-byte
// halves kept in separate 256-byte runs per nibble position: a lane's
// 16-byte read then always falls in one 256-byte run of 16 entries, which
// spans the LDS banks once, so data-dependent lookups never conflict.
// Smaller entries keep 16 per run already.
template <int EB>
__device__ __forceinline__ uint32_t lds16_index(uint32_t e, uint32_t d) {
    if constexpr (EB == 32) return (e >> 4) * 128u + (d >> 2) * 64u + (e & 15u) * 4u + (d & 3u);
    else return e * (EB / 4) + d;
}

template <int EB>
__device__ __forceinline__ void load_entry(const uint8_t *lds, uint32_t off, uint32_t (&d)[EB / 4]) {
    if constexpr (EB == 32) {
        const uint4 a = *reinterpret_cast<const uint4 *>(lds + off);
        const uint4 b = *reinterpret_cast<const uint4 *>(lds + off + 256);
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    } else if constexpr (EB == 16) {
        const uint4 a = *reinterpret_cast<const uint4 *>(lds + off);
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
    } else {
        const uint2 a = *reinterpret_cast<const uint2 *>(lds + off);
        d[0] = a.x; d[1] = a.y;
    }
}

// Fold the 8 elements of a 16-byte slice.  `tab` = the term's 64 entries
// (nibble position q major, value n minor) of EB bytes, laid out by
// lds16_index: nibble position q starts at byte 16 * EB * q, value n at 16n
// for 32-byte entries (halves 256 bytes apart), EB * n otherwise.
template <int EB>
__device__ __forceinline__ void fold_term16(uint32_t (&acc)[8][EB / 4], const uint8_t *tab, const uint32_t (&w)[4]) {
    constexpr int SH = EB == 32 ? 4 : (EB == 16 ? 4 : 3);  // log2 of the stride between values n
    constexpr uint32_t Q = 16u * EB;                         // bytes per nibble position
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t x = (w[j] >> (16 * h)) & 0xffffu;
            uint32_t(&a)[EB / 4] = acc[2 * j + h];
            if constexpr (EB == 32) {
                // Two lookups in flight at a time keep the register budget
                // of the 16-row accumulators at four waves per SIMD.
                uint32_t e0[8], e1[8];
                load_entry<EB>(tab, ((x & 15u) << SH), e0);
                load_entry<EB>(tab, Q + (((x >> 4) & 15u) << SH), e1);
#pragma unroll
                for (int d = 0; d < 8; d++) a[d] = xor3(a[d], e0[d], e1[d]);
                __builtin_amdgcn_sched_barrier(0);
                load_entry<EB>(tab, 2 * Q + (((x >> 8) & 15u) << SH), e0);
                load_entry<EB>(tab, 3 * Q + ((x >> 12) << SH), e1);
#pragma unroll
                for (int d = 0; d < 8; d++) a[d] = xor3(a[d], e0[d], e1[d]);
                __builtin_amdgcn_sched_barrier(0);
            } else {
                uint32_t e0[EB / 4], e1[EB / 4], e2[EB / 4], e3[EB / 4];
                load_entry<EB>(tab, ((x & 15u) << SH), e0);
                load_entry<EB>(tab, Q + (((x >> 4) & 15u) << SH), e1);
                load_entry<EB>(tab, 2 * Q + (((x >> 8) & 15u) << SH), e2);
                load_entry<EB>(tab, 3 * Q + ((x >> 12) << SH), e3);
#pragma unroll
                for (int d = 0; d < EB / 4; d++) a[d] = xor3(xor3(a[d], e0[d], e1[d]), e2[d], e3[d]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

}  // namespace rowsdev
}  // namespace scrs
