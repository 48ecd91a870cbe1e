// Goldilocks field (p = 2^64 - 2^32 + 1) and its cubic extension.
//
// Semantics follow pil2-stark/src/goldilocks/src/goldilocks_base_field_scalar.hpp
// (add :16-42, sub :69-103, mul :229-260) and goldilocks_cubic_extension.hpp:232-251,
// except that every value leaving gl_add/gl_sub/gl_mul is canonical (< p). The
// reference may hold non-canonical words (>= p) and compares through toU64();
// canonical outputs therefore compare equal bit for bit after the reference's toU64.
//
// Device code (gfx950) works on 32-bit halves with VOP carry chains:
//   product  a*b = p00 + 2^32 (p01 + p10) + 2^64 p11 (four v_mad_u64_u32), combined
//            into 128 bits x = (x3:x2:x1:x0) by two interleaved carry chains;
//   reduce   2^64 = 2^32 - 1 and 2^96 = -1 (mod p) give
//              x = (x1:x0) + x2*2^32 - x2 - x3 = A - B' + cA*(2^32 - 1)
//            with A = ((x1 + x2) mod 2^32 : x0), cA its carry, B' = x2 + x3 (33 bits);
//            a borrow of A - B' wraps by 2^64 = 2^32 - 1, so the value is
//            (A - B') + (cA - borrow)*(2^32 - 1), and neither correction wraps again
//            (cA = 1 forces A < p - 1; borrow = 1 forces A - B' + 2^64 > 2^64 - 2^33).
// gfx950 needs two wait states between a VALU write of an SGPR/VCC lane mask and
// a VALU read of it (carry-in or select); the inline sequences below insert them
// (s_nop) or fill them with independent instructions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL  // 2^64 mod p

__host__ __device__ __forceinline__ uint64_t gl_canon(uint64_t a) { return a >= GL_P ? a - GL_P : a; }

__host__ __device__ __forceinline__ uint64_t gl_pack(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// (hi:lo) -> canonical value mod p (host reference form of the reduction).
__host__ __device__ __forceinline__ uint64_t gl_reduce128(uint64_t hi, uint64_t lo)
{
    uint64_t hh = hi >> 32;
    uint64_t hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;      // borrow: add p back (2^64 - (2^32-1))
    uint64_t t1 = (hl << 32) - hl;  // hl * (2^32 - 1) < p
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;        // carry: 2^64 = 2^32 - 1
    return (r >= GL_P) ? r - GL_P : r;
}

#if defined(__HIP_DEVICE_COMPILE__)
// 128-bit product of two 64-bit words, as 32-bit words x0..x3.
__device__ __forceinline__ void gl_mul_words(uint64_t a, uint64_t b, uint32_t &x0, uint32_t &x1, uint32_t &x2,
                                             uint32_t &x3)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0, p01 = (uint64_t)a0 * b1, p10 = (uint64_t)a1 * b0, p11 = (uint64_t)a1 * b1;
    uint32_t y1, y2, y3;
    uint64_t c2;
    // chain 1 (vcc):  (y1,y2,y3) = [h0,l3,h3] + [l1,h1,0]
    // chain 2 (c2):   (y1,y2,y3) += [l2,h2,0]
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
    x0 = (uint32_t)p00;
    x1 = y1;
    x2 = y2;
    x3 = y3;
}

#endif

#if defined(__HIP_DEVICE_COMPILE__)
// 128-bit value -> 64-bit word congruent mod p (not necessarily canonical).
__device__ __forceinline__ uint64_t gl_reduce_words(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3)
{
    uint32_t r0, r1, t0, t1, t2;
    uint64_t ca, mp;
    asm("v_add_co_u32_e64 %[t0], %[ca], %[x1], %[x2]\n\t"
        "v_add_co_u32_e32 %[t1], vcc, %[x2], %[x3]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t2], 0, 1, vcc\n\t"
        "v_sub_co_u32_e32 %[r0], vcc, %[x0], %[t1]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e32 %[r1], vcc, %[t0], %[t2], vcc\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[mp], %[ca], vcc\n\t"
        "s_andn2_b64 vcc, vcc, %[ca]\n\t"
        "v_cndmask_b32_e64 %[t0], 0, -1, %[mp]\n\t"
        "v_cndmask_b32_e64 %[t1], 0, -1, vcc\n\t"
        "v_cndmask_b32_e64 %[t0], %[t0], 1, vcc\n\t"
        // (r1:r0) += (t1:t0); the correction cannot wrap (see header)
        "v_add_co_u32_e32 %[r0], vcc, %[r0], %[t0]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %[r1], vcc, %[r1], %[t1], vcc"
        : [r0] "=&v"(r0), [r1] "=&v"(r1), [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [ca] "=&s"(ca),
          [mp] "=&s"(mp)
        : [x0] "v"(x0), [x1] "v"(x1), [x2] "v"(x2), [x3] "v"(x3)
        : "vcc", "scc");
    return gl_pack(r0, r1);
}

// Any 64-bit word -> canonical: r >= p  <=>  r + (2^32 - 1) carries out of 64 bits.
__device__ __forceinline__ uint64_t gl_canon_dev(uint64_t r)
{
    uint32_t lo = (uint32_t)r, hi = (uint32_t)(r >> 32), t0, t1;
    asm("v_add_co_u32_e32 %[t0], vcc, -1, %[lo]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %[t1], vcc, 0, %[hi], vcc\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e32 %[lo], %[lo], %[t0], vcc\n\t"
        "v_cndmask_b32_e32 %[hi], %[hi], %[t1], vcc"
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [lo] "+v"(lo), [hi] "+v"(hi)
        :
        : "vcc", "scc");
    return gl_pack(lo, hi);
}
#endif

// a, b canonical -> canonical a + b
__host__ __device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // s = a + b (< 2p); s - p = s + (2^32 - 1) mod 2^64; take it when either sum carries
    uint32_t s0, s1, t0, t1;
    uint64_t m;
    asm("v_add_co_u32_e32 %[s0], vcc, %[a0], %[b0]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e32 %[s1], vcc, %[a1], %[b1], vcc\n\t"
        "v_add_co_u32_e64 %[t0], %[m], -1, %[s0]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t1], %[m], 0, %[s1], %[m]\n\t"
        "s_nop 1\n\t"
        "s_or_b64 vcc, vcc, %[m]\n\t"
        "v_cndmask_b32_e32 %[s0], %[s0], %[t0], vcc\n\t"
        "v_cndmask_b32_e32 %[s1], %[s1], %[t1], vcc"
        : [s0] "=&v"(s0), [s1] "=&v"(s1), [t0] "=&v"(t0), [t1] "=&v"(t1), [m] "=&s"(m)
        : [a0] "v"((uint32_t)a), [a1] "v"((uint32_t)(a >> 32)), [b0] "v"((uint32_t)b),
          [b1] "v"((uint32_t)(b >> 32))
        : "vcc", "scc");
    return gl_pack(s0, s1);
#else
    uint64_t s = a + b;
    uint64_t adj = (s < a) ? GL_EPS : 0;  // wrapped past 2^64: subtract p == add 2^32-1
    s += adj;
    return (s >= GL_P) ? s - GL_P : s;
#endif
}

// a, b canonical -> canonical a - b
__host__ __device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // d = a - b; on borrow add p, i.e. the 32-bit words (1, 2^32 - 1)
    uint32_t d0, d1, t0, t1;
    uint64_t m;
    asm("v_sub_co_u32_e32 %[d0], vcc, %[a0], %[b0]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e32 %[d1], vcc, %[a1], %[b1], vcc\n\t"
        "v_add_co_u32_e64 %[t0], %[m], 1, %[d0]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t1], %[m], -1, %[d1], %[m]\n\t"
        "v_cndmask_b32_e32 %[d0], %[d0], %[t0], vcc\n\t"
        "v_cndmask_b32_e32 %[d1], %[d1], %[t1], vcc"
        : [d0] "=&v"(d0), [d1] "=&v"(d1), [t0] "=&v"(t0), [t1] "=&v"(t1), [m] "=&s"(m)
        : [a0] "v"((uint32_t)a), [a1] "v"((uint32_t)(a >> 32)), [b0] "v"((uint32_t)b),
          [b1] "v"((uint32_t)(b >> 32))
        : "vcc", "scc");
    return gl_pack(d0, d1);
#else
    uint64_t d = a - b;
    return (a < b) ? d + GL_P : d;
#endif
}

__host__ __device__ __forceinline__ uint64_t gl_neg(uint64_t a) { return a ? GL_P - a : 0; }

// any 64-bit a, b -> canonical a * b
__host__ __device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t x0, x1, x2, x3;
    gl_mul_words(a, b, x0, x1, x2, x3);
    return gl_canon_dev(gl_reduce_words(x0, x1, x2, x3));
#else
    unsigned __int128 x = (unsigned __int128)a * b;
    return gl_reduce128((uint64_t)(x >> 64), (uint64_t)x);
#endif
}

// any 64-bit a, b -> 64-bit word congruent to a * b (not necessarily canonical);
// for products that only feed further multiplications
__host__ __device__ __forceinline__ uint64_t gl_mul_nc(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t x0, x1, x2, x3;
    gl_mul_words(a, b, x0, x1, x2, x3);
    return gl_reduce_words(x0, x1, x2, x3);
#else
    return gl_mul(a, b);
#endif
}

// canonical value of the 128-bit number (hi:lo)
__host__ __device__ __forceinline__ uint64_t gl_reduce128_canon(uint64_t lo, uint64_t hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return gl_canon_dev(gl_reduce_words((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)));
#else
    return gl_reduce128(hi, lo);
#endif
}

__host__ __device__ __forceinline__ uint64_t gl_pow(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, b);
        b = gl_mul(b, b);
        e >>= 1;
    }
    return r;
}

__host__ __device__ __forceinline__ uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

// ---------------------------------------------------------------- cubic ext --
// F_p[x]/(x^3 - x - 1), element = (c0, c1, c2) stored interleaved as 3 u64.
struct gl3 { uint64_t v[3]; };

__host__ __device__ __forceinline__ gl3 gl3_add(const gl3 &a, const gl3 &b)
{
    gl3 r;
    for (int k = 0; k < 3; k++) r.v[k] = gl_add(a.v[k], b.v[k]);
    return r;
}

// Karatsuba-style product with x^3 = x + 1 (goldilocks_cubic_extension.hpp:232-245)
__host__ __device__ __forceinline__ gl3 gl3_mul(const gl3 &a, const gl3 &b)
{
    uint64_t A = gl_mul(gl_add(a.v[0], a.v[1]), gl_add(b.v[0], b.v[1]));
    uint64_t B = gl_mul(gl_add(a.v[0], a.v[2]), gl_add(b.v[0], b.v[2]));
    uint64_t C = gl_mul(gl_add(a.v[1], a.v[2]), gl_add(b.v[1], b.v[2]));
    uint64_t D = gl_mul(a.v[0], b.v[0]);
    uint64_t E = gl_mul(a.v[1], b.v[1]);
    uint64_t F = gl_mul(a.v[2], b.v[2]);
    uint64_t G = gl_sub(D, E);
    gl3 r;
    r.v[0] = gl_sub(gl_add(C, G), F);
    r.v[1] = gl_sub(gl_sub(gl_sub(gl_add(A, C), E), E), D);
    r.v[2] = gl_sub(B, G);
    return r;
}

__host__ __device__ __forceinline__ gl3 gl3_mul_base(const gl3 &a, uint64_t b)
{
    gl3 r;
    for (int k = 0; k < 3; k++) r.v[k] = gl_mul(a.v[k], b);
    return r;
}

__host__ __device__ __forceinline__ uint32_t bitrev32(uint32_t x, uint32_t bits)
{
    if (bits == 0) return 0;
#if defined(__HIP_DEVICE_COMPILE__)
    return __brev(x) >> (32 - bits);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    x = (x >> 16) | (x << 16);
    return x >> (32 - bits);
#endif
}
