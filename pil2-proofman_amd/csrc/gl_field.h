// Goldilocks field (p = 2^64 - 2^32 + 1) and its cubic extension.
//
// Semantics follow pil2-stark/src/goldilocks/src/goldilocks_base_field_scalar.hpp
// (add :16-42, sub :69-103, mul :229-260) and goldilocks_cubic_extension.hpp:232-251,
// except that every value leaving gl_add/gl_sub/gl_mul is canonical (< p). The
// reference may hold non-canonical words (>= p) and compares through toU64();
// canonical outputs therefore compare equal bit for bit after the reference's toU64.
//
// Device code (gfx950): products are chains of v_mad_u64_u32 whose addends carry
// the partial sums, reduced with one more v_mad_u64_u32 and a short correction
// (gl_prod_mac / gl_prod_reduce*, below); additions and subtractions are 32-bit
// carry chains. gfx950 needs two wait states between a VALU write of an SGPR/VCC
// lane mask and a VALU read of it (carry-in or select); the inline sequences
// below insert them (s_nop) or fill them with independent instructions.
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
    return gl_canon_dev(gl_prod_reduce1(gl_prod_mac(a, b, 0, 0)));
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
    return gl_prod_reduce1(gl_prod_mac(a, b, 0, 0));
#else
    return gl_mul(a, b);
#endif
}

// canonical value of the 128-bit number (hi:lo)
__host__ __device__ __forceinline__ uint64_t gl_reduce128_canon(uint64_t lo, uint64_t hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    gl_prod x;
    x.lo = lo;
    x.x2 = (uint32_t)hi;
    x.x3 = (uint32_t)(hi >> 32);
    return gl_canon_dev(gl_prod_reduce1(x));
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
