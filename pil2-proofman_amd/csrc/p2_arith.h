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
// Products are the carry-form products of gl_field.h (gl_prodk3 / gl_reducek3,
// gl_mulk1): chains of v_mad_u64_u32 whose addends carry the partial sums,
// reduced with one more v_mad_u64_u32 and a short correction, left non-canonical
// here. Three independent lanes are interleaved wherever the permutation has them.
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
    uint32_t r0[3], r1[3];
    uint64_t c[3], bb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    // +c*(2^32 - 1) on a carry-out c of the high-word add: the low word loses c
    // (borrow-out B), the high word gains c & !B. Every carry mask is read at
    // least two instructions after its write.
    asm("v_add_co_u32_e64 %[r1a], %[ca], %[r1a], %[h0a]\n\t"
        "v_add_co_u32_e64 %[r1b], %[cb], %[r1b], %[h0b]\n\t"
        "v_add_co_u32_e64 %[r1c], %[cc], %[r1c], %[h0c]\n\t"
        "v_subb_co_u32_e64 %[r0a], %[ba], %[r0a], 0, %[ca]\n\t"
        "v_subb_co_u32_e64 %[r0b], %[bb], %[r0b], 0, %[cb]\n\t"
        "v_subb_co_u32_e64 %[r0c], %[bc], %[r0c], 0, %[cc]\n\t"
        "s_andn2_b64 %[ca], %[ca], %[ba]\n\t"
        "s_andn2_b64 %[cb], %[cb], %[bb]\n\t"
        "s_andn2_b64 %[cc], %[cc], %[bc]\n\t"
        "v_addc_co_u32_e64 %[r1a], %[ba], %[r1a], 0, %[ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[bb], %[r1b], 0, %[cb]\n\t"
        "v_addc_co_u32_e64 %[r1c], %[bc], %[r1c], 0, %[cc]"
        : [r0a] "+v"(r0[0]), [r0b] "+v"(r0[1]), [r0c] "+v"(r0[2]), [r1a] "+v"(r1[0]), [r1b] "+v"(r1[1]),
          [r1c] "+v"(r1[2]), [ca] "=&s"(c[0]), [cb] "=&s"(c[1]), [cc] "=&s"(c[2]), [ba] "=&s"(bb[0]),
          [bb] "=&s"(bb[1]), [bc] "=&s"(bb[2])
        : [h0a] "v"((uint32_t)a[0].hi), [h0b] "v"((uint32_t)a[1].hi), [h0c] "v"((uint32_t)a[2].hi)
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(r0[k], r1[k]);
}

// One halves-form fold (hz_fold3's steps, hazard gaps filled with s_nop).
__device__ __forceinline__ uint64_t hz_fold(hz a)
{
    uint64_t r, d, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(d) : "v"((uint32_t)(a.hi >> 32)), "v"(a.lo));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32);
    uint64_t b;
    asm("v_add_co_u32_e64 %[r1], %[c], %[r1], %[h0]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r0], %[b], %[r0], 0, %[c]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[c], %[c], %[b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[r1], %[b], %[r1], 0, %[c]"
        : [r0] "+v"(r0), [r1] "+v"(r1), [c] "=&s"(c), [b] "=&s"(b)
        : [h0] "v"((uint32_t)a.hi)
        : "scc");
    return gl_pack(r0, r1);
}

// a*b mod p as a 64-bit word (any 64-bit a, b)
__device__ __forceinline__ uint64_t p2_mul(uint64_t a, uint64_t b) { return gl_mulk1(a, b, 0, 0); }

// a*b + s mod p as a 64-bit word (s = s_hi:s_lo)
__device__ __forceinline__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    return gl_mulk1(a, b, s_lo, s_hi);
}

// p2_mul_add with a wave-uniform b
__device__ __forceinline__ uint64_t p2_mul_add_u(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    return gl_mulk1<true>(a, b, s_lo, s_hi);
}

// out[k] = a[k]*b[k] mod p for three lanes (any 64-bit words)
__device__ __forceinline__ void p2_mul3(const uint64_t *a, const uint64_t *b, uint64_t *out)
{
    gl_prodk x[3];
    gl_prodk3(a, b, 0, 0, x);
    gl_reducek3(x, out);
}

// out[k] = a[k]*b[k] + s mod p for three lanes (s = s_hi:s_lo shared; b wave-uniform)
__device__ __forceinline__ void p2_mul_add3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi,
                                            uint64_t *out)
{
    gl_prodk x[3];
    gl_prodk3<true>(a, b, s_lo, s_hi, x);
    gl_reducek3(x, out);
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
__device__ uint64_t p2_mul_add_u(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi);

#endif
