// Non-canonical Goldilocks arithmetic for the Poseidon2 permutation (gfx950).
//
// Inside the permutation every value only has to be congruent mod p; it is
// canonicalised once, when a digest leaves the kernel. Two forms are used:
//   u64 word  any 64-bit value v (meaning v mod p);
//   lz        three 32-bit words (l0, l1, c) meaning l0 + 2^32 l1 + 2^64 c,
//             with a small carry count c: the result of additions whose
//             2^64-overflows are counted instead of reduced.
// Since 2^64 = 2^32 - 1 (mod p), an lz value folds back to a word with one
// v_mad_u64_u32: (l1:l0) + c*(2^32 - 1), plus one conditional +(2^32 - 1) when
// that sum carries out (the carried remainder is < c*2^32, so it cannot carry
// again).
//
// Additions run as 32-bit carry chains. gfx950 needs two wait states between
// a VALU write of a carry mask and the VALU instruction reading it; the
// batched adders below interleave three or four independent chains (carries in
// vcc and SGPR pairs) so that every read is at least two instructions after
// its write, and no s_nop is needed.
//
// Products: a*b + s for 64-bit a, b, s is assembled from four v_mad_u64_u32
// partial products, with s entering as the 32-bit addends of a0*b0 (low half)
// and a0*b1 (high half): (2^32-1)^2 + (2^32-1) < 2^64, so neither overflows.
// The 128-bit result goes through gl_reduce_words (gl_field.h).
#pragma once
#include "gl_field.h"

struct lz {
    uint32_t l0, l1, c;
};

__host__ __device__ __forceinline__ lz lz_of(uint64_t v) { return {(uint32_t)v, (uint32_t)(v >> 32), 0u}; }

// a * 2^s for s = 1, 2 (no carries: the count absorbs the shifted-out bits)
template <int S>
__host__ __device__ __forceinline__ lz lz_shl(lz a)
{
    lz r;
    r.c = (a.c << S) | (a.l1 >> (32 - S));
    r.l1 = (a.l1 << S) | (a.l0 >> (32 - S));
    r.l0 = a.l0 << S;
    return r;
}

#if defined(__HIP_DEVICE_COMPILE__)

// r[k] = a[k] + b[k] for k < N (N = 1..4), chains interleaved.
template <int N>
__device__ __forceinline__ void lz_add_n(lz *r, const lz *a, const lz *b);

template <>
__device__ __forceinline__ void lz_add_n<1>(lz *r, const lz *a, const lz *b)
{
    asm("v_add_co_u32_e32 %0, vcc, %3, %6\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %1, vcc, %4, %7, vcc\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %2, vcc, %5, %8, vcc"
        : "=&v"(r[0].l0), "=&v"(r[0].l1), "=&v"(r[0].c)
        : "v"(a[0].l0), "v"(a[0].l1), "v"(a[0].c), "v"(b[0].l0), "v"(b[0].l1), "v"(b[0].c)
        : "vcc");
}

template <>
__device__ __forceinline__ void lz_add_n<2>(lz *r, const lz *a, const lz *b)
{
    uint64_t k;
    asm("v_add_co_u32_e32 %0, vcc, %7, %10\n\t"
        "v_add_co_u32_e64 %3, %[k], %13, %16\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e32 %1, vcc, %8, %11, vcc\n\t"
        "v_addc_co_u32_e64 %4, %[k], %14, %17, %[k]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e32 %2, vcc, %9, %12, vcc\n\t"
        "v_addc_co_u32_e64 %5, %[k], %15, %18, %[k]"
        : "=&v"(r[0].l0), "=&v"(r[0].l1), "=&v"(r[0].c), "=&v"(r[1].l0), "=&v"(r[1].l1), "=&v"(r[1].c),
          [k] "=&s"(k)
        : "v"(a[0].l0), "v"(a[0].l1), "v"(a[0].c), "v"(b[0].l0), "v"(b[0].l1), "v"(b[0].c), "v"(a[1].l0),
          "v"(a[1].l1), "v"(a[1].c), "v"(b[1].l0), "v"(b[1].l1), "v"(b[1].c)
        : "vcc");
}

template <>
__device__ __forceinline__ void lz_add_n<3>(lz *r, const lz *a, const lz *b)
{
    uint64_t k1, k2;
    asm("v_add_co_u32_e32 %0, vcc, %11, %14\n\t"
        "v_add_co_u32_e64 %3, %[k1], %17, %20\n\t"
        "v_add_co_u32_e64 %6, %[k2], %23, %26\n\t"
        "v_addc_co_u32_e32 %1, vcc, %12, %15, vcc\n\t"
        "v_addc_co_u32_e64 %4, %[k1], %18, %21, %[k1]\n\t"
        "v_addc_co_u32_e64 %7, %[k2], %24, %27, %[k2]\n\t"
        "v_addc_co_u32_e32 %2, vcc, %13, %16, vcc\n\t"
        "v_addc_co_u32_e64 %5, %[k1], %19, %22, %[k1]\n\t"
        "v_addc_co_u32_e64 %8, %[k2], %25, %28, %[k2]"
        : "=&v"(r[0].l0), "=&v"(r[0].l1), "=&v"(r[0].c), "=&v"(r[1].l0), "=&v"(r[1].l1), "=&v"(r[1].c),
          "=&v"(r[2].l0), "=&v"(r[2].l1), "=&v"(r[2].c), [k1] "=&s"(k1), [k2] "=&s"(k2)
        : "v"(a[0].l0), "v"(a[0].l1), "v"(a[0].c), "v"(b[0].l0), "v"(b[0].l1), "v"(b[0].c), "v"(a[1].l0),
          "v"(a[1].l1), "v"(a[1].c), "v"(b[1].l0), "v"(b[1].l1), "v"(b[1].c), "v"(a[2].l0), "v"(a[2].l1),
          "v"(a[2].c), "v"(b[2].l0), "v"(b[2].l1), "v"(b[2].c)
        : "vcc");
}

template <>
__device__ __forceinline__ void lz_add_n<4>(lz *r, const lz *a, const lz *b)
{
    lz_add_n<2>(r, a, b);
    lz_add_n<2>(r + 2, a + 2, b + 2);
}

// r[k] = a[k] + b[k] for k < n, in interleaved groups of three.
template <int N>
__device__ __forceinline__ void lz_add_all(lz *r, const lz *a, const lz *b)
{
    int k = 0;
#pragma unroll
    for (; k + 3 <= N; k += 3) lz_add_n<3>(r + k, a + k, b + k);
    if constexpr (N % 3 == 1) lz_add_n<1>(r + N - 1, a + N - 1, b + N - 1);
    if constexpr (N % 3 == 2) lz_add_n<2>(r + N - 2, a + N - 2, b + N - 2);
}

// lz -> congruent 64-bit word (c < 2^31)
__device__ __forceinline__ uint64_t lz_fold(lz a)
{
    uint64_t lo = gl_pack(a.l0, a.l1), r, cc;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=s"(cc) : "v"(a.c), "v"(lo));
    uint64_t re = r + GL_EPS;  // the carried remainder is < 2^63, so this cannot wrap
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32);
    asm("s_nop 1\n\t"
        "v_cndmask_b32_e64 %0, %0, %2, %4\n\t"
        "v_cndmask_b32_e64 %1, %1, %3, %4"
        : "+v"(r0), "+v"(r1)
        : "v"((uint32_t)re), "v"((uint32_t)(re >> 32)), "s"(cc));
    return gl_pack(r0, r1);
}

// 128-bit (x3:x2:x1:x0) from the four partial products of a 64x64 product,
// p = p00 + 2^32 (p01 + p10) + 2^64 p11 (each pij < 2^64).
__device__ __forceinline__ uint64_t gl_combine_reduce(uint64_t p00, uint64_t p01, uint64_t p10, uint64_t p11)
{
    uint32_t y1, y2, y3;
    uint64_t c2;
    asm("v_add_co_u32_e32 %[y1], vcc, %[h0], %[l1]\n\t"
        "v_add_co_u32_e64 %[y1], %[c2], %[y1], %[l2]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e32 %[y2], vcc, %[l3], %[h1], vcc\n\t"
        "v_addc_co_u32_e64 %[y2], %[c2], %[y2], %[h2], %[c2]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e32 %[y3], vcc, 0, %[h3], vcc\n\t"
        "v_addc_co_u32_e64 %[y3], %[c2], %[y3], 0, %[c2]"
        : [y1] "=&v"(y1), [y2] "=&v"(y2), [y3] "=&v"(y3), [c2] "=&s"(c2)
        : [h0] "v"((uint32_t)(p00 >> 32)), [l1] "v"((uint32_t)p01), [h1] "v"((uint32_t)(p01 >> 32)),
          [l2] "v"((uint32_t)p10), [h2] "v"((uint32_t)(p10 >> 32)), [l3] "v"((uint32_t)p11),
          [h3] "v"((uint32_t)(p11 >> 32))
        : "vcc", "scc");
    return gl_reduce_words((uint32_t)p00, y1, y2, y3);
}

// a*b mod p as a 64-bit word (any 64-bit a, b)
__device__ __forceinline__ uint64_t p2_mul(uint64_t a, uint64_t b)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    return gl_combine_reduce((uint64_t)a0 * b0, (uint64_t)a0 * b1, (uint64_t)a1 * b0, (uint64_t)a1 * b1);
}

// a*b + s mod p as a 64-bit word; s enters through the multiply-add addends
__device__ __forceinline__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    return gl_combine_reduce((uint64_t)a0 * b0 + s_lo, (uint64_t)a0 * b1 + s_hi, (uint64_t)a1 * b0,
                             (uint64_t)a1 * b1);
}

#else  // host pass: declarations only (the kernels are compiled for the device)

template <int N> __device__ void lz_add_n(lz *r, const lz *a, const lz *b);
template <int N> __device__ void lz_add_all(lz *r, const lz *a, const lz *b);
__device__ uint64_t lz_fold(lz a);
__device__ uint64_t p2_mul(uint64_t a, uint64_t b);
__device__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi);

#endif
