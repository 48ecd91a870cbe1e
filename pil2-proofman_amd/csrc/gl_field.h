// Goldilocks field (p = 2^64 - 2^32 + 1) and its cubic extension, device side.
//
// Semantics follow pil2-stark/src/goldilocks/src/goldilocks_base_field_scalar.hpp
// (add :16-42, sub :69-103, mul :229-260) and goldilocks_cubic_extension.hpp:232-251,
// except that every value leaving a function here is canonical (< p). The reference
// may hold non-canonical words (>= p) and compares through toU64(); canonical
// outputs therefore compare equal bit for bit after the reference's toU64.
//
// Multiplication is a 64x64->128 integer product on the VALU followed by the
// Goldilocks reduction that uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL  // 2^64 mod p

__host__ __device__ __forceinline__ uint64_t gl_canon(uint64_t a) { return a >= GL_P ? a - GL_P : a; }

// a, b canonical -> canonical
__host__ __device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    // overflow past 2^64: subtract p == add 2^32-1 modulo 2^64
    uint64_t adj = (s < a) ? GL_EPS : 0;
    s += adj;
    return (s >= GL_P) ? s - GL_P : s;
}

__host__ __device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b)
{
    uint64_t d = a - b;
    return (a < b) ? d + GL_P : d;
}

__host__ __device__ __forceinline__ uint64_t gl_neg(uint64_t a) { return a ? GL_P - a : 0; }

// (hi:lo) -> canonical value mod p.
__host__ __device__ __forceinline__ uint64_t gl_reduce128(uint64_t hi, uint64_t lo)
{
    uint64_t hh = hi >> 32;
    uint64_t hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;           // borrow: add p back (2^64 - (2^32-1))
    uint64_t t1 = (hl << 32) - hl;       // hl * (2^32 - 1) < p
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;             // carry: 2^64 = 2^32 - 1
    return (r >= GL_P) ? r - GL_P : r;
}

// Same reduction without the final canonicalisation: result < 2^64 and
// congruent; used inside Poseidon2 where inputs to mul may be any 64-bit word.
__host__ __device__ __forceinline__ uint64_t gl_reduce128_lazy(uint64_t hi, uint64_t lo)
{
    uint64_t hh = hi >> 32;
    uint64_t hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;
    uint64_t t1 = (hl << 32) - hl;
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;
    return r;
}

__host__ __device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return gl_reduce128(__umul64hi(a, b), a * b);
#else
    unsigned __int128 x = (unsigned __int128)a * b;
    return gl_reduce128((uint64_t)(x >> 64), (uint64_t)x);
#endif
}

__host__ __device__ __forceinline__ uint64_t gl_mul_lazy(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return gl_reduce128_lazy(__umul64hi(a, b), a * b);
#else
    unsigned __int128 x = (unsigned __int128)a * b;
    return gl_reduce128_lazy((uint64_t)(x >> 64), (uint64_t)x);
#endif
}

// Lazy add: a any 64-bit word, b any 64-bit word -> congruent 64-bit word.
__host__ __device__ __forceinline__ uint64_t gl_add_lazy(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    if (s < a) {                 // wrapped: add 2^32-1; cannot wrap twice unless s >= 2^64-2^32+1
        s += GL_EPS;
        if (s < GL_EPS) s += GL_EPS;
    }
    return s;
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
