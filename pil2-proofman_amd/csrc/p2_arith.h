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

// ---- products in carry form --------------------------------------------
// a*b + s (a, b any 64-bit words, s = s_hi:s_lo) as four v_mad_u64_u32 whose
// addends carry the partial sums:
//   p = a0*b0 + s_lo                       x0 = p mod 2^32
//   t = a0*b1 + (p >> 32) + s_hi           (< 2^64)
//   u = a1*b0 + t, carry-out k (2^96)      x1 = u mod 2^32
//   h = a1*b1 + (u >> 32)                  (x3:x2) = h
// so a*b + s = (x1:x0) + h*2^64 + k*2^96, and x3 + k < 2^32 (the product is
// below 2^128). Taking t whole as the addend of u leaves the top carry k in an
// SGPR lane mask instead of a third addition into h; the reduction subtracts it
// as the borrow-in of its x3 subtraction.
struct p2_prodk {
    uint64_t lo;  // (x1:x0)
    uint64_t h;   // (x3:x2)
    uint64_t k;   // lane mask: the carry k
};

// Three lanes; the carry masks are written by one asm block, so every later
// read is at least two instructions after its write.
__device__ __forceinline__ void p2_prodk3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi,
                                          p2_prodk *x)
{
    uint64_t p[3], t[3], u[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uint32_t a0 = (uint32_t)a[i], b0 = (uint32_t)b[i], b1 = (uint32_t)(b[i] >> 32);
        p[i] = (uint64_t)a0 * b0 + s_lo;
        t[i] = (uint64_t)a0 * b1 + ((p[i] >> 32) + (uint64_t)s_hi);
    }
    asm("v_mad_u64_u32 %0, %3, %6, %9, %12\n\t"
        "v_mad_u64_u32 %1, %4, %7, %10, %13\n\t"
        "v_mad_u64_u32 %2, %5, %8, %11, %14"
        : "=&v"(u[0]), "=&v"(u[1]), "=&v"(u[2]), "=&s"(x[0].k), "=&s"(x[1].k), "=&s"(x[2].k)
        : "v"((uint32_t)(a[0] >> 32)), "v"((uint32_t)(a[1] >> 32)), "v"((uint32_t)(a[2] >> 32)), "v"((uint32_t)b[0]),
          "v"((uint32_t)b[1]), "v"((uint32_t)b[2]), "v"(t[0]), "v"(t[1]), "v"(t[2]));
#pragma unroll
    for (int i = 0; i < 3; i++) {
        x[i].h = (uint64_t)(uint32_t)(a[i] >> 32) * (uint32_t)(b[i] >> 32) + (u[i] >> 32);
        x[i].lo = gl_pack((uint32_t)p[i], (uint32_t)u[i]);
    }
}

// Reduction of three carry-form products to congruent 64-bit words. With
// 2^64 = 2^32 - 1 = e and 2^96 = -1 (mod p):
//   r = (x1:x0) + x2*e, carry-out c          (one v_mad_u64_u32)
//   r = r - x3 - k, borrow-out b             (k is the borrow-in)
//   r = r + (c - b)*e                        (never wraps: c = 1 gives r <= 2^64 - 2^33
//                                             before the subtraction; b = 1 with c = 0
//                                             gives r >= 2^64 - 2^32 + 1 after it)
// The correction is one add-with-carry per half: with m = c & !b and n = b & !c,
// lo = r0 + (m ? 2^32 - 1 : 0) + n and hi = r1 + (n ? 2^32 - 1 : 0) + carry, which
// adds e for m and 2^64 - e (i.e. -e mod 2^64) for n.
__device__ __forceinline__ void p2_reducek3(const p2_prodk *x, uint64_t *out)
{
    uint64_t r[3], c[3];
    asm("v_mad_u64_u32 %0, %3, %6, -1, %9\n\t"
        "v_mad_u64_u32 %1, %4, %7, -1, %10\n\t"
        "v_mad_u64_u32 %2, %5, %8, -1, %11"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&s"(c[0]), "=&s"(c[1]), "=&s"(c[2])
        : "v"((uint32_t)x[0].h), "v"((uint32_t)x[1].h), "v"((uint32_t)x[2].h), "v"(x[0].lo), "v"(x[1].lo),
          "v"(x[2].lo));
    uint32_t r0[3], r1[3], tm[3], tn[3];
    uint64_t b[3], m[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r0[i] = (uint32_t)r[i];
        r1[i] = (uint32_t)(r[i] >> 32);
    }
    asm("v_subb_co_u32_e64 %[r0a], %[ba], %[r0a], %[x3a], %[ka]\n\t"
        "v_subb_co_u32_e64 %[r0b], %[bb], %[r0b], %[x3b], %[kb]\n\t"
        "v_subb_co_u32_e64 %[r0c], %[bc], %[r0c], %[x3c], %[kc]\n\t"
        "v_subb_co_u32_e64 %[r1a], %[ba], %[r1a], 0, %[ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[bb], %[r1b], 0, %[bb]\n\t"
        "v_subb_co_u32_e64 %[r1c], %[bc], %[r1c], 0, %[bc]\n\t"
        "s_andn2_b64 %[ma], %[ca], %[ba]\n\t"
        "s_andn2_b64 %[mb], %[cb], %[bb]\n\t"
        "s_andn2_b64 %[mc], %[cc], %[bc]\n\t"
        "s_andn2_b64 %[ba], %[ba], %[ca]\n\t"
        "s_andn2_b64 %[bb], %[bb], %[cb]\n\t"
        "s_andn2_b64 %[bc], %[bc], %[cc]\n\t"
        "v_cndmask_b32_e64 %[tma], 0, -1, %[ma]\n\t"
        "v_cndmask_b32_e64 %[tmb], 0, -1, %[mb]\n\t"
        "v_cndmask_b32_e64 %[tmc], 0, -1, %[mc]\n\t"
        "v_cndmask_b32_e64 %[tna], 0, -1, %[ba]\n\t"
        "v_cndmask_b32_e64 %[tnb], 0, -1, %[bb]\n\t"
        "v_cndmask_b32_e64 %[tnc], 0, -1, %[bc]\n\t"
        "v_addc_co_u32_e64 %[r0a], %[ca], %[r0a], %[tma], %[ba]\n\t"
        "v_addc_co_u32_e64 %[r0b], %[cb], %[r0b], %[tmb], %[bb]\n\t"
        "v_addc_co_u32_e64 %[r0c], %[cc], %[r0c], %[tmc], %[bc]\n\t"
        "v_addc_co_u32_e64 %[r1a], %[ca], %[r1a], %[tna], %[ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[cb], %[r1b], %[tnb], %[cb]\n\t"
        "v_addc_co_u32_e64 %[r1c], %[cc], %[r1c], %[tnc], %[cc]"
        : [r0a] "+v"(r0[0]), [r0b] "+v"(r0[1]), [r0c] "+v"(r0[2]), [r1a] "+v"(r1[0]), [r1b] "+v"(r1[1]),
          [r1c] "+v"(r1[2]), [tma] "=&v"(tm[0]), [tmb] "=&v"(tm[1]), [tmc] "=&v"(tm[2]), [tna] "=&v"(tn[0]),
          [tnb] "=&v"(tn[1]), [tnc] "=&v"(tn[2]), [ba] "=&s"(b[0]), [bb] "=&s"(b[1]), [bc] "=&s"(b[2]),
          [ma] "=&s"(m[0]), [mb] "=&s"(m[1]), [mc] "=&s"(m[2]), [ca] "+s"(c[0]), [cb] "+s"(c[1]), [cc] "+s"(c[2])
        : [x3a] "v"((uint32_t)(x[0].h >> 32)), [x3b] "v"((uint32_t)(x[1].h >> 32)),
          [x3c] "v"((uint32_t)(x[2].h >> 32)), [ka] "s"(x[0].k), [kb] "s"(x[1].k), [kc] "s"(x[2].k)
        : "scc");
#pragma unroll
    for (int i = 0; i < 3; i++) out[i] = gl_pack(r0[i], r1[i]);
}

// One lane of p2_prodk3 / p2_reducek3, hazard gaps filled with s_nop.
__device__ __forceinline__ uint64_t p2_mulk1(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p = (uint64_t)a0 * b0 + s_lo;
    uint64_t t = (uint64_t)a0 * b1 + ((p >> 32) + (uint64_t)s_hi);
    uint64_t u, k;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=&v"(u), "=&s"(k) : "v"(a1), "v"(b0), "v"(t));
    uint64_t h = (uint64_t)a1 * b1 + (u >> 32);
    uint64_t lo = gl_pack((uint32_t)p, (uint32_t)u), r, c, bm, m;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"((uint32_t)h), "v"(lo));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), tm, tn;
    asm("s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r0], %[b], %[r0], %[x3], %[k]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r1], %[b], %[r1], 0, %[b]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[m], %[c], %[b]\n\t"
        "s_andn2_b64 %[b], %[b], %[c]\n\t"
        "v_cndmask_b32_e64 %[tm], 0, -1, %[m]\n\t"
        "v_cndmask_b32_e64 %[tn], 0, -1, %[b]\n\t"
        "v_addc_co_u32_e64 %[r0], %[c], %[r0], %[tm], %[b]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[r1], %[c], %[r1], %[tn], %[c]"
        : [r0] "+v"(r0), [r1] "+v"(r1), [tm] "=&v"(tm), [tn] "=&v"(tn), [b] "=&s"(bm), [m] "=&s"(m), [c] "+s"(c)
        : [x3] "v"((uint32_t)(h >> 32)), [k] "s"(k)
        : "scc");
    return gl_pack(r0, r1);
}

// a*b mod p as a 64-bit word (any 64-bit a, b)
__device__ __forceinline__ uint64_t p2_mul(uint64_t a, uint64_t b) { return p2_mulk1(a, b, 0, 0); }

// a*b + s mod p as a 64-bit word (s = s_hi:s_lo)
__device__ __forceinline__ uint64_t p2_mul_add(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    return p2_mulk1(a, b, s_lo, s_hi);
}

// out[k] = a[k]*b[k] mod p for three lanes (any 64-bit words)
__device__ __forceinline__ void p2_mul3(const uint64_t *a, const uint64_t *b, uint64_t *out)
{
    p2_prodk x[3];
    p2_prodk3(a, b, 0, 0, x);
    p2_reducek3(x, out);
}

// out[k] = a[k]*b[k] + s mod p for three lanes (s = s_hi:s_lo shared)
__device__ __forceinline__ void p2_mul_add3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi,
                                            uint64_t *out)
{
    p2_prodk x[3];
    p2_prodk3(a, b, s_lo, s_hi, x);
    p2_reducek3(x, out);
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
