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

// r[k] = a[k] + rc[k] for three lanes of round constants (wave-uniform): the
// low halves of the constants are read from SGPRs directly (the carry-in steps
// need VGPR operands, so only the high halves are copied).
__device__ __forceinline__ void lz_add_rc3(lz *r, const lz *a, const uint64_t *rc)
{
    uint64_t k1, k2;
    asm("v_add_co_u32_e32 %0, vcc, %11, %14\n\t"
        "v_add_co_u32_e64 %3, %[k1], %12, %16\n\t"
        "v_add_co_u32_e64 %6, %[k2], %13, %18\n\t"
        "v_addc_co_u32_e32 %1, vcc, %15, %20, vcc\n\t"
        "v_addc_co_u32_e64 %4, %[k1], %17, %21, %[k1]\n\t"
        "v_addc_co_u32_e64 %7, %[k2], %19, %22, %[k2]\n\t"
        "v_addc_co_u32_e32 %2, vcc, 0, %23, vcc\n\t"
        "v_addc_co_u32_e64 %5, %[k1], 0, %24, %[k1]\n\t"
        "v_addc_co_u32_e64 %8, %[k2], 0, %25, %[k2]"
        : "=&v"(r[0].l0), "=&v"(r[0].l1), "=&v"(r[0].c), "=&v"(r[1].l0), "=&v"(r[1].l1), "=&v"(r[1].c),
          "=&v"(r[2].l0), "=&v"(r[2].l1), "=&v"(r[2].c), [k1] "=&s"(k1), [k2] "=&s"(k2)
        : "s"((uint32_t)rc[0]), "s"((uint32_t)rc[1]), "s"((uint32_t)rc[2]), "v"(a[0].l0), "v"(a[0].l1),
          "v"(a[1].l0), "v"(a[1].l1), "v"(a[2].l0), "v"(a[2].l1), "v"((uint32_t)(rc[0] >> 32)),
          "v"((uint32_t)(rc[1] >> 32)), "v"((uint32_t)(rc[2] >> 32)), "v"(a[0].c), "v"(a[1].c), "v"(a[2].c)
        : "vcc");
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

// Three independent 128-bit products (partial products p00..p11 per lane, as in
// gl_combine_reduce) assembled and reduced in lockstep: every instruction of a
// lane is three instructions after the previous one of the same lane, which
// covers the two wait states a carry-mask read needs after its VALU write, so
// the sequence has no s_nop. Per lane: (x3:x2:x1:x0) by two carry chains, then
// the reduction of gl_reduce_words (x = (A:x0) - B' + (cA - borrow)*(2^32 - 1)).
__device__ __forceinline__ void gl_combine_reduce3(const uint64_t *p00, const uint64_t *p01, const uint64_t *p10,
                                                   const uint64_t *p11, uint64_t *out)
{
    uint32_t l0[3], h0[3], y1[3], y2[3], y3[3];
    uint64_t mA[3], mB[3], mC[3], mD[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        l0[k] = (uint32_t)p00[k];
        h0[k] = (uint32_t)(p00[k] >> 32);
    }
    asm(
        "v_add_co_u32_e64 %[y1a], %[Aa], %[h0a], %[l1a]\n\t"
        "v_add_co_u32_e64 %[y1b], %[Ab], %[h0b], %[l1b]\n\t"
        "v_add_co_u32_e64 %[y1c], %[Ac], %[h0c], %[l1c]\n\t"
        "v_add_co_u32_e64 %[y1a], %[Ba], %[y1a], %[l2a]\n\t"
        "v_add_co_u32_e64 %[y1b], %[Bb], %[y1b], %[l2b]\n\t"
        "v_add_co_u32_e64 %[y1c], %[Bc], %[y1c], %[l2c]\n\t"
        "v_addc_co_u32_e64 %[y2a], %[Aa], %[l3a], %[h1a], %[Aa]\n\t"
        "v_addc_co_u32_e64 %[y2b], %[Ab], %[l3b], %[h1b], %[Ab]\n\t"
        "v_addc_co_u32_e64 %[y2c], %[Ac], %[l3c], %[h1c], %[Ac]\n\t"
        "v_addc_co_u32_e64 %[y2a], %[Ba], %[y2a], %[h2a], %[Ba]\n\t"
        "v_addc_co_u32_e64 %[y2b], %[Bb], %[y2b], %[h2b], %[Bb]\n\t"
        "v_addc_co_u32_e64 %[y2c], %[Bc], %[y2c], %[h2c], %[Bc]\n\t"
        "v_addc_co_u32_e64 %[y3a], %[Aa], 0, %[h3a], %[Aa]\n\t"
        "v_addc_co_u32_e64 %[y3b], %[Ab], 0, %[h3b], %[Ab]\n\t"
        "v_addc_co_u32_e64 %[y3c], %[Ac], 0, %[h3c], %[Ac]\n\t"
        "v_addc_co_u32_e64 %[y3a], %[Ba], %[y3a], 0, %[Ba]\n\t"
        "v_addc_co_u32_e64 %[y3b], %[Bb], %[y3b], 0, %[Bb]\n\t"
        "v_addc_co_u32_e64 %[y3c], %[Bc], %[y3c], 0, %[Bc]\n\t"
        "v_add_co_u32_e64 %[h0a], %[Aa], %[y1a], %[y2a]\n\t"
        "v_add_co_u32_e64 %[h0b], %[Ab], %[y1b], %[y2b]\n\t"
        "v_add_co_u32_e64 %[h0c], %[Ac], %[y1c], %[y2c]\n\t"
        "v_add_co_u32_e64 %[y1a], %[Ba], %[y2a], %[y3a]\n\t"
        "v_add_co_u32_e64 %[y1b], %[Bb], %[y2b], %[y3b]\n\t"
        "v_add_co_u32_e64 %[y1c], %[Bc], %[y2c], %[y3c]\n\t"
        "v_cndmask_b32_e64 %[y2a], 0, 1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[y2b], 0, 1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[y2c], 0, 1, %[Bc]\n\t"
        "v_sub_co_u32_e64 %[l0a], %[Ca], %[l0a], %[y1a]\n\t"
        "v_sub_co_u32_e64 %[l0b], %[Cb], %[l0b], %[y1b]\n\t"
        "v_sub_co_u32_e64 %[l0c], %[Cc], %[l0c], %[y1c]\n\t"
        "v_subb_co_u32_e64 %[h0a], %[Ca], %[h0a], %[y2a], %[Ca]\n\t"
        "v_subb_co_u32_e64 %[h0b], %[Cb], %[h0b], %[y2b], %[Cb]\n\t"
        "v_subb_co_u32_e64 %[h0c], %[Cc], %[h0c], %[y2c], %[Cc]\n\t"
        "s_andn2_b64 %[Ba], %[Aa], %[Ca]\n\t"
        "s_andn2_b64 %[Bb], %[Ab], %[Cb]\n\t"
        "s_andn2_b64 %[Bc], %[Ac], %[Cc]\n\t"
        "s_andn2_b64 %[Da], %[Ca], %[Aa]\n\t"
        "s_andn2_b64 %[Db], %[Cb], %[Ab]\n\t"
        "s_andn2_b64 %[Dc], %[Cc], %[Ac]\n\t"
        "v_cndmask_b32_e64 %[y1a], 0, -1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[y1b], 0, -1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[y1c], 0, -1, %[Bc]\n\t"
        "v_cndmask_b32_e64 %[y3a], 0, -1, %[Da]\n\t"
        "v_cndmask_b32_e64 %[y3b], 0, -1, %[Db]\n\t"
        "v_cndmask_b32_e64 %[y3c], 0, -1, %[Dc]\n\t"
        "v_cndmask_b32_e64 %[y1a], %[y1a], 1, %[Da]\n\t"
        "v_cndmask_b32_e64 %[y1b], %[y1b], 1, %[Db]\n\t"
        "v_cndmask_b32_e64 %[y1c], %[y1c], 1, %[Dc]\n\t"
        "v_add_co_u32_e64 %[l0a], %[Aa], %[l0a], %[y1a]\n\t"
        "v_add_co_u32_e64 %[l0b], %[Ab], %[l0b], %[y1b]\n\t"
        "v_add_co_u32_e64 %[l0c], %[Ac], %[l0c], %[y1c]\n\t"
        "v_addc_co_u32_e64 %[h0a], %[Aa], %[h0a], %[y3a], %[Aa]\n\t"
        "v_addc_co_u32_e64 %[h0b], %[Ab], %[h0b], %[y3b], %[Ab]\n\t"
        "v_addc_co_u32_e64 %[h0c], %[Ac], %[h0c], %[y3c], %[Ac]"
        : [l0a] "+v"(l0[0]), [h0a] "+v"(h0[0]), [l0b] "+v"(l0[1]), [h0b] "+v"(h0[1]), [l0c] "+v"(l0[2]),
          [h0c] "+v"(h0[2]), [y1a] "=&v"(y1[0]), [y2a] "=&v"(y2[0]), [y3a] "=&v"(y3[0]), [y1b] "=&v"(y1[1]),
          [y2b] "=&v"(y2[1]), [y3b] "=&v"(y3[1]), [y1c] "=&v"(y1[2]), [y2c] "=&v"(y2[2]), [y3c] "=&v"(y3[2]),
          [Aa] "=&s"(mA[0]), [Ba] "=&s"(mB[0]), [Ca] "=&s"(mC[0]), [Da] "=&s"(mD[0]), [Ab] "=&s"(mA[1]),
          [Bb] "=&s"(mB[1]), [Cb] "=&s"(mC[1]), [Db] "=&s"(mD[1]), [Ac] "=&s"(mA[2]), [Bc] "=&s"(mB[2]),
          [Cc] "=&s"(mC[2]), [Dc] "=&s"(mD[2])
        : [l1a] "v"((uint32_t)p01[0]), [h1a] "v"((uint32_t)(p01[0] >> 32)), [l2a] "v"((uint32_t)p10[0]),
          [h2a] "v"((uint32_t)(p10[0] >> 32)), [l3a] "v"((uint32_t)p11[0]), [h3a] "v"((uint32_t)(p11[0] >> 32)),
          [l1b] "v"((uint32_t)p01[1]), [h1b] "v"((uint32_t)(p01[1] >> 32)), [l2b] "v"((uint32_t)p10[1]),
          [h2b] "v"((uint32_t)(p10[1] >> 32)), [l3b] "v"((uint32_t)p11[1]), [h3b] "v"((uint32_t)(p11[1] >> 32)),
          [l1c] "v"((uint32_t)p01[2]), [h1c] "v"((uint32_t)(p01[2] >> 32)), [l2c] "v"((uint32_t)p10[2]),
          [h2c] "v"((uint32_t)(p10[2] >> 32)), [l3c] "v"((uint32_t)p11[2]), [h3c] "v"((uint32_t)(p11[2] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(l0[k], h0[k]);
}

// Three folds (lz_fold) with the carry fix-ups interleaved.
__device__ __forceinline__ void lz_fold3(const lz *a, uint64_t *out)
{
    uint64_t r[3], cc[3];
    asm("v_mad_u64_u32 %0, %3, %6, -1, %9\n\t"
        "v_mad_u64_u32 %1, %4, %7, -1, %10\n\t"
        "v_mad_u64_u32 %2, %5, %8, -1, %11"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&s"(cc[0]), "=&s"(cc[1]), "=&s"(cc[2])
        : "v"(a[0].c), "v"(a[1].c), "v"(a[2].c), "v"(gl_pack(a[0].l0, a[0].l1)), "v"(gl_pack(a[1].l0, a[1].l1)),
          "v"(gl_pack(a[2].l0, a[2].l1)));
    uint32_t r0[3], r1[3], t[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    // the mads above are at least two instructions before each mask read below
    asm("v_cndmask_b32_e64 %6, 0, -1, %9\n\t"
        "v_cndmask_b32_e64 %7, 0, -1, %10\n\t"
        "v_cndmask_b32_e64 %8, 0, -1, %11\n\t"
        "v_add_co_u32_e64 %0, %9, %0, %6\n\t"
        "v_add_co_u32_e64 %1, %10, %1, %7\n\t"
        "v_add_co_u32_e64 %2, %11, %2, %8\n\t"
        "v_addc_co_u32_e64 %3, %9, 0, %3, %9\n\t"
        "v_addc_co_u32_e64 %4, %10, 0, %4, %10\n\t"
        "v_addc_co_u32_e64 %5, %11, 0, %5, %11"
        : "+v"(r0[0]), "+v"(r0[1]), "+v"(r0[2]), "+v"(r1[0]), "+v"(r1[1]), "+v"(r1[2]), "=&v"(t[0]), "=&v"(t[1]),
          "=&v"(t[2]), "+s"(cc[0]), "+s"(cc[1]), "+s"(cc[2]));
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(r0[k], r1[k]);
}


// ---- multiply-accumulate products (no carry chains) ----------------------
// a*b for 64-bit a = a1:a0, b = b1:b0 as a chain of v_mad_u64_u32 whose 64-bit
// addends absorb the previous partial product, so no step can overflow
// ((2^32-1)^2 + 2*(2^32-1) = 2^64 - 1):
//   p = a0*b0 + s_lo                  x0 = p mod 2^32
//   t = a0*b1 + (p >> 32) + s_hi
//   u = a1*b0 + (t mod 2^32)          x1 = u mod 2^32
//   (x3:x2) = a1*b1 + (t >> 32) + (u >> 32)
// (s = s_hi:s_lo is an optional addend of the whole product.) Reduction then
// uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod p):
//   v = (x1:x0) + x2*(2^32 - 1) - x3,
// computed as r = mad(x2, 2^32 - 1, x1:x0) with carry-out c, then r - x3 with
// borrow b, then + (c - b)*(2^32 - 1): c = 1 forces r <= 2^64 - 2^33 and b = 1
// with c = 0 forces r - x3 + 2^64 > 2^64 - 2^32, so the correction never wraps.
struct gl_prod {
    uint64_t lo;     // (x1:x0)
    uint32_t x2, x3;
};

__device__ __forceinline__ gl_prod gl_prod_mac(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p = (uint64_t)a0 * b0 + s_lo;
    uint64_t t = (uint64_t)a0 * b1 + ((p >> 32) + (uint64_t)s_hi);
    uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
    uint64_t h = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
    gl_prod r;
    r.lo = gl_pack((uint32_t)p, (uint32_t)u);
    r.x2 = (uint32_t)h;
    r.x3 = (uint32_t)(h >> 32);
    return r;
}

// Reduce three products in lockstep (three instructions between dependent
// carry-mask writes and reads: no s_nop).
__device__ __forceinline__ void gl_prod_reduce3(const gl_prod *x, uint64_t *out)
{
    uint64_t r[3], c[3];
    asm("v_mad_u64_u32 %0, %3, %6, -1, %9\n\t"
        "v_mad_u64_u32 %1, %4, %7, -1, %10\n\t"
        "v_mad_u64_u32 %2, %5, %8, -1, %11"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&s"(c[0]), "=&s"(c[1]), "=&s"(c[2])
        : "v"(x[0].x2), "v"(x[1].x2), "v"(x[2].x2), "v"(x[0].lo), "v"(x[1].lo), "v"(x[2].lo));
    uint32_t r0[3], r1[3], t0[3], t1[3];
    uint64_t b[3], m[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    asm("v_sub_co_u32_e64 %[r0a], %[ba], %[r0a], %[x3a]\n\t"
        "v_sub_co_u32_e64 %[r0b], %[bb], %[r0b], %[x3b]\n\t"
        "v_sub_co_u32_e64 %[r0c], %[bc], %[r0c], %[x3c]\n\t"
        "v_subb_co_u32_e64 %[r1a], %[ba], %[r1a], 0, %[ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[bb], %[r1b], 0, %[bb]\n\t"
        "v_subb_co_u32_e64 %[r1c], %[bc], %[r1c], 0, %[bc]\n\t"
        "s_andn2_b64 %[ma], %[ca], %[ba]\n\t"
        "s_andn2_b64 %[mb], %[cb], %[bb]\n\t"
        "s_andn2_b64 %[mc], %[cc], %[bc]\n\t"
        "s_andn2_b64 %[ba], %[ba], %[ca]\n\t"
        "s_andn2_b64 %[bb], %[bb], %[cb]\n\t"
        "s_andn2_b64 %[bc], %[bc], %[cc]\n\t"
        "v_cndmask_b32_e64 %[t0a], 0, -1, %[ma]\n\t"
        "v_cndmask_b32_e64 %[t0b], 0, -1, %[mb]\n\t"
        "v_cndmask_b32_e64 %[t0c], 0, -1, %[mc]\n\t"
        "v_cndmask_b32_e64 %[t1a], 0, -1, %[ba]\n\t"
        "v_cndmask_b32_e64 %[t1b], 0, -1, %[bb]\n\t"
        "v_cndmask_b32_e64 %[t1c], 0, -1, %[bc]\n\t"
        "v_cndmask_b32_e64 %[t0a], %[t0a], 1, %[ba]\n\t"
        "v_cndmask_b32_e64 %[t0b], %[t0b], 1, %[bb]\n\t"
        "v_cndmask_b32_e64 %[t0c], %[t0c], 1, %[bc]\n\t"
        "v_add_co_u32_e64 %[r0a], %[ca], %[r0a], %[t0a]\n\t"
        "v_add_co_u32_e64 %[r0b], %[cb], %[r0b], %[t0b]\n\t"
        "v_add_co_u32_e64 %[r0c], %[cc], %[r0c], %[t0c]\n\t"
        "v_addc_co_u32_e64 %[r1a], %[ca], %[r1a], %[t1a], %[ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[cb], %[r1b], %[t1b], %[cb]\n\t"
        "v_addc_co_u32_e64 %[r1c], %[cc], %[r1c], %[t1c], %[cc]"
        : [r0a] "+v"(r0[0]), [r0b] "+v"(r0[1]), [r0c] "+v"(r0[2]), [r1a] "+v"(r1[0]), [r1b] "+v"(r1[1]),
          [r1c] "+v"(r1[2]), [t0a] "=&v"(t0[0]), [t0b] "=&v"(t0[1]), [t0c] "=&v"(t0[2]), [t1a] "=&v"(t1[0]),
          [t1b] "=&v"(t1[1]), [t1c] "=&v"(t1[2]), [ba] "=&s"(b[0]), [bb] "=&s"(b[1]), [bc] "=&s"(b[2]),
          [ma] "=&s"(m[0]), [mb] "=&s"(m[1]), [mc] "=&s"(m[2]), [ca] "+s"(c[0]), [cb] "+s"(c[1]), [cc] "+s"(c[2])
        : [x3a] "v"(x[0].x3), [x3b] "v"(x[1].x3), [x3c] "v"(x[2].x3)
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(r0[k], r1[k]);
}

// Single product reduction (hazard gaps filled with s_nop).
__device__ __forceinline__ uint64_t gl_prod_reduce1(gl_prod x)
{
    uint64_t r, c, b, m;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"(x.x2), "v"(x.lo));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), t0, t1;
    asm("v_sub_co_u32_e64 %[r0], %[b], %[r0], %[x3]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r1], %[b], %[r1], 0, %[b]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[m], %[c], %[b]\n\t"
        "s_andn2_b64 %[b], %[b], %[c]\n\t"
        "v_cndmask_b32_e64 %[t0], 0, -1, %[m]\n\t"
        "v_cndmask_b32_e64 %[t1], 0, -1, %[b]\n\t"
        "v_cndmask_b32_e64 %[t0], %[t0], 1, %[b]\n\t"
        "v_add_co_u32_e64 %[r0], %[c], %[r0], %[t0]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[r1], %[c], %[r1], %[t1], %[c]"
        : [r0] "+v"(r0), [r1] "+v"(r1), [t0] "=&v"(t0), [t1] "=&v"(t1), [b] "=&s"(b), [m] "=&s"(m), [c] "+s"(c)
        : [x3] "v"(x.x3)
        : "scc");
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
template <int N> __device__ void lz_add_all(lz *r, const lz *a, const lz *b);
__device__ void lz_add_rc3(lz *r, const lz *a, const uint64_t *rc);
__device__ void p2_mul3(const uint64_t *a, const uint64_t *b, uint64_t *out);
__device__ void p2_mul_add3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi, uint64_t *out);
__device__ void lz_fold3(const lz *a, uint64_t *out);
__device__ uint64_t lz_fold(lz a);
__device__ uint64_t p2_mul(uint64_t a, uint64_t b);
__device__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi);

#endif
