// Non-canonical Goldilocks arithmetic for the Poseidon2 permutation (gfx950).
//
// Inside the permutation every value only has to be congruent mod p; it is
// canonicalised once, when a digest leaves the kernel. Three forms are used:
//   u64 word  any 64-bit value v (meaning v mod p);
//   hz        halves form: two 64-bit accumulators (lo, hi) meaning lo + 2^32 hi,
//             each below 2^41. The linear layers sum the 32-bit halves of words
//             with small coefficients in it, with no carry chains at all;
//   lz        three 32-bit words (l0, l1, c) meaning l0 + 2^32 l1 + 2^64 c with a
//             small carry count c (the partial rounds' round-constant addition).
// Since 2^64 = 2^32 - 1 (mod p), both fold back to a word with one
// v_mad_u64_u32 and a conditional +(2^32 - 1) on a carry-out.
//
// gfx950 needs two wait states between a VALU write of a carry mask and the
// VALU instruction reading it; the three-lane helpers below interleave three
// independent chains (carries in SGPR pairs) so that every read is at least two
// instructions after its write, and the single-lane ones fill the gaps with
// s_nop.
//
// Products are chains of v_mad_u64_u32 whose addends carry the partial sums
// (no carry chains at all), reduced with one more v_mad_u64_u32 and a short
// correction; gfx950 issues the carry and mask instructions at single rate
// while plain moves dual-issue, so this form minimises the single-rate count.
// Three independent lanes are interleaved wherever the permutation has them.
#pragma once
#include "gl_field.h"

struct lz {
    uint32_t l0, l1, c;
};

// Halves form: two 64-bit accumulators (lo, hi) meaning lo + 2^32 hi, each below
// 2^41. Linear combinations with small coefficients of the 32-bit halves of
// 64-bit words stay far below 2^64, so they need no carry chains.
struct hz {
    uint64_t lo, hi;
};

__host__ __device__ __forceinline__ lz lz_of(uint64_t v) { return {(uint32_t)v, (uint32_t)(v >> 32), 0u}; }

#if defined(__HIP_DEVICE_COMPILE__)

// r[k] = a[k] + b[k] for k < N (the permutation uses N = 1).
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

// lz -> congruent 64-bit word (c < 2^31)
__device__ __forceinline__ uint64_t lz_fold(lz a)
{
    uint64_t lo = gl_pack(a.l0, a.l1), r, cc;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=s"(cc) : "v"(a.c), "v"(lo));
    // on carry-out add 2^32 - 1 (the carried remainder is < c*2^32, so this cannot wrap)
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), t;
    asm("s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t], 0, -1, %[cc]\n\t"
        "v_add_co_u32_e32 %[r0], vcc, %[r0], %[t]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %[r1], vcc, 0, %[r1], vcc"
        : [r0] "+v"(r0), [r1] "+v"(r1), [t] "=&v"(t)
        : [cc] "s"(cc)
        : "vcc");
    return gl_pack(r0, r1);
}

// Three halves-form values (lo, hi < 2^41) to congruent 64-bit words. With
// hi = h0 + 2^32 h1 and 2^64 = 2^32 - 1 (mod p): r = lo + h1 (2^32 - 1) is
// below 2^42 (one v_mad_u64_u32, no carry-out); then h0 is added to r's high
// word, and a carry-out of that add (2^64) becomes + (2^32 - 1), which cannot
// carry again since the wrapped value is below 2^42.
__device__ __forceinline__ void hz_fold3(const hz *a, uint64_t *out)
{
    uint64_t r[3], d[3];
    asm("v_mad_u64_u32 %0, %3, %6, -1, %9\n\t"
        "v_mad_u64_u32 %1, %4, %7, -1, %10\n\t"
        "v_mad_u64_u32 %2, %5, %8, -1, %11"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&s"(d[0]), "=&s"(d[1]), "=&s"(d[2])
        : "v"((uint32_t)(a[0].hi >> 32)), "v"((uint32_t)(a[1].hi >> 32)), "v"((uint32_t)(a[2].hi >> 32)),
          "v"(a[0].lo), "v"(a[1].lo), "v"(a[2].lo));
    uint32_t r0[3], r1[3], t[3];
    uint64_t c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    // every carry mask is read at least two instructions after its write
    asm("v_add_co_u32_e64 %3, %9, %3, %12\n\t"
        "v_add_co_u32_e64 %4, %10, %4, %13\n\t"
        "v_add_co_u32_e64 %5, %11, %5, %14\n\t"
        "v_cndmask_b32_e64 %6, 0, -1, %9\n\t"
        "v_cndmask_b32_e64 %7, 0, -1, %10\n\t"
        "v_cndmask_b32_e64 %8, 0, -1, %11\n\t"
        "v_add_co_u32_e64 %0, %9, %0, %6\n\t"
        "v_add_co_u32_e64 %1, %10, %1, %7\n\t"
        "v_add_co_u32_e64 %2, %11, %2, %8\n\t"
        "v_addc_co_u32_e64 %3, %9, 0, %3, %9\n\t"
        "v_addc_co_u32_e64 %4, %10, 0, %4, %10\n\t"
        "v_addc_co_u32_e64 %5, %11, 0, %5, %11"
        : "+v"(r0[0]), "+v"(r0[1]), "+v"(r0[2]), "+v"(r1[0]), "+v"(r1[1]), "+v"(r1[2]), "=&v"(t[0]), "=&v"(t[1]),
          "=&v"(t[2]), "=&s"(c[0]), "=&s"(c[1]), "=&s"(c[2])
        : "v"((uint32_t)a[0].hi), "v"((uint32_t)a[1].hi), "v"((uint32_t)a[2].hi));
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(r0[k], r1[k]);
}

// One halves-form fold (hz_fold3's steps, hazard gaps filled with s_nop).
__device__ __forceinline__ uint64_t hz_fold(hz a)
{
    uint64_t r, d, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(d) : "v"((uint32_t)(a.hi >> 32)), "v"(a.lo));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), t;
    asm("v_add_co_u32_e64 %[r1], %[c], %[r1], %[h0]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t], 0, -1, %[c]\n\t"
        "v_add_co_u32_e64 %[r0], %[c], %[r0], %[t]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[r1], %[c], 0, %[r1], %[c]"
        : [r0] "+v"(r0), [r1] "+v"(r1), [t] "=&v"(t), [c] "=&s"(c)
        : [h0] "v"((uint32_t)a.hi));
    return gl_pack(r0, r1);
}

// a*b mod p as a 64-bit word (any 64-bit a, b)
__device__ __forceinline__ uint64_t p2_mul(uint64_t a, uint64_t b) { return gl_prod_reduce1(gl_prod_mac(a, b, 0, 0)); }

// a*b + s mod p as a 64-bit word (s = s_hi:s_lo)
__device__ __forceinline__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    return gl_prod_reduce1(gl_prod_mac(a, b, s_lo, s_hi));
}

// out[k] = a[k]*b[k] mod p for three lanes (any 64-bit words)
__device__ __forceinline__ void p2_mul3(const uint64_t *a, const uint64_t *b, uint64_t *out)
{
    gl_prod x[3];
#pragma unroll
    for (int k = 0; k < 3; k++) x[k] = gl_prod_mac(a[k], b[k], 0, 0);
    gl_prod_reduce3(x, out);
}

// out[k] = a[k]*b[k] + s mod p for three lanes (s = s_hi:s_lo shared)
__device__ __forceinline__ void p2_mul_add3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi,
                                            uint64_t *out)
{
    gl_prod x[3];
#pragma unroll
    for (int k = 0; k < 3; k++) x[k] = gl_prod_mac(a[k], b[k], s_lo, s_hi);
    gl_prod_reduce3(x, out);
}

#else  // host pass: declarations only (the kernels are compiled for the device)

template <int N> __device__ void lz_add_n(lz *r, const lz *a, const lz *b);
__device__ void p2_mul3(const uint64_t *a, const uint64_t *b, uint64_t *out);
__device__ void p2_mul_add3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi, uint64_t *out);
__device__ void hz_fold3(const hz *a, uint64_t *out);
__device__ uint64_t hz_fold(hz a);
__device__ uint64_t lz_fold(lz a);
__device__ uint64_t p2_mul(uint64_t a, uint64_t b);
__device__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi);

#endif
