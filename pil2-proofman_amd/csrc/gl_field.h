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
struct gl_prodk {
    uint64_t lo;  // (x1:x0)
    uint64_t h;   // (x3:x2)
    uint64_t k;   // lane mask: the carry k
};

// Three lanes; the carry masks are written by one asm block, so every later
// read is at least two instructions after its write. BU: b is wave-uniform (held
// in SGPRs, e.g. a constant of the permutation), so the asm takes it as a scalar
// operand instead of a copy in a VGPR.
template <bool BU = false>
__device__ __forceinline__ void gl_prodk3(const uint64_t *a, const uint64_t *b, uint32_t s_lo, uint32_t s_hi,
                                          gl_prodk *x)
{
    uint64_t p[3], t[3], u[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uint32_t a0 = (uint32_t)a[i], b0 = (uint32_t)b[i], b1 = (uint32_t)(b[i] >> 32);
        p[i] = (uint64_t)a0 * b0 + s_lo;
        t[i] = (uint64_t)a0 * b1 + ((p[i] >> 32) + (uint64_t)s_hi);
    }
    if constexpr (BU)
        asm("v_mad_u64_u32 %0, %3, %6, %9, %12\n\t"
            "v_mad_u64_u32 %1, %4, %7, %10, %13\n\t"
            "v_mad_u64_u32 %2, %5, %8, %11, %14"
            : "=&v"(u[0]), "=&v"(u[1]), "=&v"(u[2]), "=&s"(x[0].k), "=&s"(x[1].k), "=&s"(x[2].k)
            : "v"((uint32_t)(a[0] >> 32)), "v"((uint32_t)(a[1] >> 32)), "v"((uint32_t)(a[2] >> 32)),
              "s"((uint32_t)b[0]), "s"((uint32_t)b[1]), "s"((uint32_t)b[2]), "v"(t[0]), "v"(t[1]), "v"(t[2]));
    else
        asm("v_mad_u64_u32 %0, %3, %6, %9, %12\n\t"
            "v_mad_u64_u32 %1, %4, %7, %10, %13\n\t"
            "v_mad_u64_u32 %2, %5, %8, %11, %14"
            : "=&v"(u[0]), "=&v"(u[1]), "=&v"(u[2]), "=&s"(x[0].k), "=&s"(x[1].k), "=&s"(x[2].k)
            : "v"((uint32_t)(a[0] >> 32)), "v"((uint32_t)(a[1] >> 32)), "v"((uint32_t)(a[2] >> 32)),
              "v"((uint32_t)b[0]), "v"((uint32_t)b[1]), "v"((uint32_t)b[2]), "v"(t[0]), "v"(t[1]), "v"(t[2]));
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
__device__ __forceinline__ void gl_reducek3(const gl_prodk *x, uint64_t *out)
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

// One lane of gl_prodk3 / gl_reducek3, hazard gaps filled with s_nop.
template <bool BU = false>
__device__ __forceinline__ uint64_t gl_mulk1(uint64_t a, uint64_t b, uint32_t s_lo, uint32_t s_hi)
{
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p = (uint64_t)a0 * b0 + s_lo;
    uint64_t t = (uint64_t)a0 * b1 + ((p >> 32) + (uint64_t)s_hi);
    uint64_t u, k;
    if constexpr (BU)
        asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=&v"(u), "=&s"(k) : "v"(a1), "s"(b0), "v"(t));
    else
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


// Three arbitrary 64-bit words to canonical form in place: r >= p exactly when
// r + (2^32 - 1) carries out of 64 bits, and then that sum is r - p.
__device__ __forceinline__ void gl_canon3(uint64_t *v)
{
    uint32_t s0[3], s1[3], t0[3], t1[3];
    uint64_t K[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s0[k] = (uint32_t)v[k];
        s1[k] = (uint32_t)(v[k] >> 32);
    }
    asm("v_add_co_u32_e64 %[t0a], %[Ka], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Kb], -1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Kc], -1, %[s0c]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ka], 0, %[s1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Kb], 0, %[s1b], %[Kb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Kc], 0, %[s1c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]"
        : [s0a] "+v"(s0[0]), [s1a] "+v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [s0b] "+v"(s0[1]), [s1b] "+v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]), [Kb] "=&s"(K[1]),
          [s0c] "+v"(s0[2]), [s1c] "+v"(s1[2]), [t0c] "=&v"(t0[2]), [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2])
        :
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] = gl_pack(s0[k], s1[k]);
}

template <int N> __device__ __forceinline__ void gl_add_n(uint64_t *out, const uint64_t *a, const uint64_t *b);
template <int N> __device__ __forceinline__ void gl_sub_n(uint64_t *out, const uint64_t *a, const uint64_t *b);
template <int N> __device__ __forceinline__ void gl_prod_reduce_canon_n(uint64_t *out, const gl_prod *x);

// ---- lane-interleaved canonical operations (N = 1..4 independent lanes) ----
// Every instruction of a lane is N instructions after the previous one of the
// same lane, so for N >= 3 carry-mask reads need no s_nop; N = 1, 2 pad with
// s_nop to the two wait states gfx950 needs.

template <> __device__ __forceinline__ void gl_add_n<1>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[1], s1[1], t0[1], t1[1];
    uint64_t K[1], M[1];
    asm(
        "v_add_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], -1, %[s0a]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], 0, %[s1a], %[Ma]\n\t"
        "s_nop 1\n\t"
        "s_or_b64 %[Ka], %[Ka], %[Ma]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 1; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_sub_n<1>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[1], s1[1], t0[1], t1[1];
    uint64_t K[1], M[1];
    asm(
        "v_sub_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], 1, %[s0a]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], -1, %[s1a], %[Ma]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 1; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_prod_reduce_canon_n<1>(uint64_t *out, const gl_prod *x)
{
    uint64_t r[1], c[1], B[1], P[1], K[1];
    uint32_t r0[1], r1[1], t0[1], t1[1];
#pragma unroll
    for (int k = 0; k < 1; k++) {
        asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r[k]), "=&s"(c[k]) : "v"(x[k].x2), "v"(x[k].lo));
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    asm(
        "v_sub_co_u32_e64 %[r0a], %[Ba], %[r0a], %[x3a]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r1a], %[Ba], %[r1a], 0, %[Ba]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[Pa], %[Ca], %[Ba]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[Ba], %[Ba], %[Ca]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t0a], 0, -1, %[Pa]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t1a], 0, -1, %[Ba]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[t0a], %[t0a], 1, %[Ba]\n\t"
        "s_nop 1\n\t"
        "v_add_co_u32_e64 %[r0a], %[Ca], %[r0a], %[t0a]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[r1a], %[Ca], %[r1a], %[t1a], %[Ca]"
        : [r0a] "+v"(r0[0]), [r1a] "+v"(r1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ba] "=&s"(B[0]),
          [Pa] "=&s"(P[0]), [Ca] "+s"(c[0])
        : [x3a] "v"(x[0].x3)
        : "scc");
    uint32_t *s0 = r0, *s1 = r1;
    asm(
        "v_add_co_u32_e64 %[t0a], %[Ka], -1, %[s0a]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ka], 0, %[s1a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]"
        : [s0a] "+v"(s0[0]), [s1a] "+v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0])
        :
        : "scc");
#pragma unroll
    for (int k = 0; k < 1; k++) out[k] = gl_pack(r0[k], r1[k]);
}

template <> __device__ __forceinline__ void gl_add_n<2>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[2], s1[2], t0[2], t1[2];
    uint64_t K[2], M[2];
    asm(
        "v_add_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_add_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], -1, %[s0b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], 0, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], 0, %[s1b], %[Mb]\n\t"
        "s_nop 0\n\t"
        "s_or_b64 %[Ka], %[Ka], %[Ma]\n\t"
        "s_or_b64 %[Kb], %[Kb], %[Mb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 2; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_sub_n<2>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[2], s1[2], t0[2], t1[2];
    uint64_t K[2], M[2];
    asm(
        "v_sub_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_sub_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_subb_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], 1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], 1, %[s0b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], -1, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], -1, %[s1b], %[Mb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 2; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_prod_reduce_canon_n<2>(uint64_t *out, const gl_prod *x)
{
    uint64_t r[2], c[2], B[2], P[2], K[2];
    uint32_t r0[2], r1[2], t0[2], t1[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r[k]), "=&s"(c[k]) : "v"(x[k].x2), "v"(x[k].lo));
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    asm(
        "v_sub_co_u32_e64 %[r0a], %[Ba], %[r0a], %[x3a]\n\t"
        "v_sub_co_u32_e64 %[r0b], %[Bb], %[r0b], %[x3b]\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32_e64 %[r1a], %[Ba], %[r1a], 0, %[Ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[Bb], %[r1b], 0, %[Bb]\n\t"
        "s_nop 0\n\t"
        "s_andn2_b64 %[Pa], %[Ca], %[Ba]\n\t"
        "s_andn2_b64 %[Pb], %[Cb], %[Bb]\n\t"
        "s_nop 0\n\t"
        "s_andn2_b64 %[Ba], %[Ba], %[Ca]\n\t"
        "s_andn2_b64 %[Bb], %[Bb], %[Cb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[t0a], 0, -1, %[Pa]\n\t"
        "v_cndmask_b32_e64 %[t0b], 0, -1, %[Pb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[t1a], 0, -1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t1b], 0, -1, %[Bb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[t0a], %[t0a], 1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t0b], %[t0b], 1, %[Bb]\n\t"
        "s_nop 0\n\t"
        "v_add_co_u32_e64 %[r0a], %[Ca], %[r0a], %[t0a]\n\t"
        "v_add_co_u32_e64 %[r0b], %[Cb], %[r0b], %[t0b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[r1a], %[Ca], %[r1a], %[t1a], %[Ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[Cb], %[r1b], %[t1b], %[Cb]"
        : [r0a] "+v"(r0[0]), [r1a] "+v"(r1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ba] "=&s"(B[0]),
          [Pa] "=&s"(P[0]), [Ca] "+s"(c[0]), [r0b] "+v"(r0[1]), [r1b] "+v"(r1[1]), [t0b] "=&v"(t0[1]),
          [t1b] "=&v"(t1[1]), [Bb] "=&s"(B[1]), [Pb] "=&s"(P[1]), [Cb] "+s"(c[1])
        : [x3a] "v"(x[0].x3), [x3b] "v"(x[1].x3)
        : "scc");
    uint32_t *s0 = r0, *s1 = r1;
    asm(
        "v_add_co_u32_e64 %[t0a], %[Ka], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Kb], -1, %[s0b]\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ka], 0, %[s1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Kb], 0, %[s1b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "s_nop 0\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]"
        : [s0a] "+v"(s0[0]), [s1a] "+v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [s0b] "+v"(s0[1]), [s1b] "+v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]), [Kb] "=&s"(K[1])
        :
        : "scc");
#pragma unroll
    for (int k = 0; k < 2; k++) out[k] = gl_pack(r0[k], r1[k]);
}

template <> __device__ __forceinline__ void gl_add_n<3>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[3], s1[3], t0[3], t1[3];
    uint64_t K[3], M[3];
    asm(
        "v_add_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_add_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "v_add_co_u32_e64 %[s0c], %[Kc], %[a0c], %[b0c]\n\t"
        "v_addc_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "v_addc_co_u32_e64 %[s1c], %[Kc], %[a1c], %[b1c], %[Kc]\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], -1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Mc], -1, %[s0c]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], 0, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], 0, %[s1b], %[Mb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Mc], 0, %[s1c], %[Mc]\n\t"
        "s_or_b64 %[Ka], %[Ka], %[Ma]\n\t"
        "s_or_b64 %[Kb], %[Kb], %[Mb]\n\t"
        "s_or_b64 %[Kc], %[Kc], %[Mc]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1]), [s0c] "=&v"(s0[2]), [s1c] "=&v"(s1[2]), [t0c] "=&v"(t0[2]),
          [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2]), [Mc] "=&s"(M[2])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32)), [a0c] "v"((uint32_t)a[2]),
          [a1c] "v"((uint32_t)(a[2] >> 32)), [b0c] "v"((uint32_t)b[2]), [b1c] "v"((uint32_t)(b[2] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_sub_n<3>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[3], s1[3], t0[3], t1[3];
    uint64_t K[3], M[3];
    asm(
        "v_sub_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_sub_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "v_sub_co_u32_e64 %[s0c], %[Kc], %[a0c], %[b0c]\n\t"
        "v_subb_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_subb_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "v_subb_co_u32_e64 %[s1c], %[Kc], %[a1c], %[b1c], %[Kc]\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], 1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], 1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Mc], 1, %[s0c]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], -1, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], -1, %[s1b], %[Mb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Mc], -1, %[s1c], %[Mc]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1]), [s0c] "=&v"(s0[2]), [s1c] "=&v"(s1[2]), [t0c] "=&v"(t0[2]),
          [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2]), [Mc] "=&s"(M[2])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32)), [a0c] "v"((uint32_t)a[2]),
          [a1c] "v"((uint32_t)(a[2] >> 32)), [b0c] "v"((uint32_t)b[2]), [b1c] "v"((uint32_t)(b[2] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_prod_reduce_canon_n<3>(uint64_t *out, const gl_prod *x)
{
    uint64_t r[3], c[3], B[3], P[3], K[3];
    uint32_t r0[3], r1[3], t0[3], t1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r[k]), "=&s"(c[k]) : "v"(x[k].x2), "v"(x[k].lo));
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    asm(
        "v_sub_co_u32_e64 %[r0a], %[Ba], %[r0a], %[x3a]\n\t"
        "v_sub_co_u32_e64 %[r0b], %[Bb], %[r0b], %[x3b]\n\t"
        "v_sub_co_u32_e64 %[r0c], %[Bc], %[r0c], %[x3c]\n\t"
        "v_subb_co_u32_e64 %[r1a], %[Ba], %[r1a], 0, %[Ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[Bb], %[r1b], 0, %[Bb]\n\t"
        "v_subb_co_u32_e64 %[r1c], %[Bc], %[r1c], 0, %[Bc]\n\t"
        "s_andn2_b64 %[Pa], %[Ca], %[Ba]\n\t"
        "s_andn2_b64 %[Pb], %[Cb], %[Bb]\n\t"
        "s_andn2_b64 %[Pc], %[Cc], %[Bc]\n\t"
        "s_andn2_b64 %[Ba], %[Ba], %[Ca]\n\t"
        "s_andn2_b64 %[Bb], %[Bb], %[Cb]\n\t"
        "s_andn2_b64 %[Bc], %[Bc], %[Cc]\n\t"
        "v_cndmask_b32_e64 %[t0a], 0, -1, %[Pa]\n\t"
        "v_cndmask_b32_e64 %[t0b], 0, -1, %[Pb]\n\t"
        "v_cndmask_b32_e64 %[t0c], 0, -1, %[Pc]\n\t"
        "v_cndmask_b32_e64 %[t1a], 0, -1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t1b], 0, -1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[t1c], 0, -1, %[Bc]\n\t"
        "v_cndmask_b32_e64 %[t0a], %[t0a], 1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t0b], %[t0b], 1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[t0c], %[t0c], 1, %[Bc]\n\t"
        "v_add_co_u32_e64 %[r0a], %[Ca], %[r0a], %[t0a]\n\t"
        "v_add_co_u32_e64 %[r0b], %[Cb], %[r0b], %[t0b]\n\t"
        "v_add_co_u32_e64 %[r0c], %[Cc], %[r0c], %[t0c]\n\t"
        "v_addc_co_u32_e64 %[r1a], %[Ca], %[r1a], %[t1a], %[Ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[Cb], %[r1b], %[t1b], %[Cb]\n\t"
        "v_addc_co_u32_e64 %[r1c], %[Cc], %[r1c], %[t1c], %[Cc]"
        : [r0a] "+v"(r0[0]), [r1a] "+v"(r1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ba] "=&s"(B[0]),
          [Pa] "=&s"(P[0]), [Ca] "+s"(c[0]), [r0b] "+v"(r0[1]), [r1b] "+v"(r1[1]), [t0b] "=&v"(t0[1]),
          [t1b] "=&v"(t1[1]), [Bb] "=&s"(B[1]), [Pb] "=&s"(P[1]), [Cb] "+s"(c[1]), [r0c] "+v"(r0[2]),
          [r1c] "+v"(r1[2]), [t0c] "=&v"(t0[2]), [t1c] "=&v"(t1[2]), [Bc] "=&s"(B[2]), [Pc] "=&s"(P[2]),
          [Cc] "+s"(c[2])
        : [x3a] "v"(x[0].x3), [x3b] "v"(x[1].x3), [x3c] "v"(x[2].x3)
        : "scc");
    uint32_t *s0 = r0, *s1 = r1;
    asm(
        "v_add_co_u32_e64 %[t0a], %[Ka], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Kb], -1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Kc], -1, %[s0c]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ka], 0, %[s1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Kb], 0, %[s1b], %[Kb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Kc], 0, %[s1c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]"
        : [s0a] "+v"(s0[0]), [s1a] "+v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [s0b] "+v"(s0[1]), [s1b] "+v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]), [Kb] "=&s"(K[1]),
          [s0c] "+v"(s0[2]), [s1c] "+v"(s1[2]), [t0c] "=&v"(t0[2]), [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2])
        :
        : "scc");
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_pack(r0[k], r1[k]);
}

template <> __device__ __forceinline__ void gl_add_n<4>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[4], s1[4], t0[4], t1[4];
    uint64_t K[4], M[4];
    asm(
        "v_add_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_add_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "v_add_co_u32_e64 %[s0c], %[Kc], %[a0c], %[b0c]\n\t"
        "v_add_co_u32_e64 %[s0d], %[Kd], %[a0d], %[b0d]\n\t"
        "v_addc_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "v_addc_co_u32_e64 %[s1c], %[Kc], %[a1c], %[b1c], %[Kc]\n\t"
        "v_addc_co_u32_e64 %[s1d], %[Kd], %[a1d], %[b1d], %[Kd]\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], -1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Mc], -1, %[s0c]\n\t"
        "v_add_co_u32_e64 %[t0d], %[Md], -1, %[s0d]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], 0, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], 0, %[s1b], %[Mb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Mc], 0, %[s1c], %[Mc]\n\t"
        "v_addc_co_u32_e64 %[t1d], %[Md], 0, %[s1d], %[Md]\n\t"
        "s_or_b64 %[Ka], %[Ka], %[Ma]\n\t"
        "s_or_b64 %[Kb], %[Kb], %[Mb]\n\t"
        "s_or_b64 %[Kc], %[Kc], %[Mc]\n\t"
        "s_or_b64 %[Kd], %[Kd], %[Md]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s0d], %[s0d], %[t0d], %[Kd]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1d], %[s1d], %[t1d], %[Kd]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1]), [s0c] "=&v"(s0[2]), [s1c] "=&v"(s1[2]), [t0c] "=&v"(t0[2]),
          [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2]), [Mc] "=&s"(M[2]), [s0d] "=&v"(s0[3]), [s1d] "=&v"(s1[3]),
          [t0d] "=&v"(t0[3]), [t1d] "=&v"(t1[3]), [Kd] "=&s"(K[3]), [Md] "=&s"(M[3])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32)), [a0c] "v"((uint32_t)a[2]),
          [a1c] "v"((uint32_t)(a[2] >> 32)), [b0c] "v"((uint32_t)b[2]), [b1c] "v"((uint32_t)(b[2] >> 32)),
          [a0d] "v"((uint32_t)a[3]), [a1d] "v"((uint32_t)(a[3] >> 32)), [b0d] "v"((uint32_t)b[3]),
          [b1d] "v"((uint32_t)(b[3] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_sub_n<4>(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    uint32_t s0[4], s1[4], t0[4], t1[4];
    uint64_t K[4], M[4];
    asm(
        "v_sub_co_u32_e64 %[s0a], %[Ka], %[a0a], %[b0a]\n\t"
        "v_sub_co_u32_e64 %[s0b], %[Kb], %[a0b], %[b0b]\n\t"
        "v_sub_co_u32_e64 %[s0c], %[Kc], %[a0c], %[b0c]\n\t"
        "v_sub_co_u32_e64 %[s0d], %[Kd], %[a0d], %[b0d]\n\t"
        "v_subb_co_u32_e64 %[s1a], %[Ka], %[a1a], %[b1a], %[Ka]\n\t"
        "v_subb_co_u32_e64 %[s1b], %[Kb], %[a1b], %[b1b], %[Kb]\n\t"
        "v_subb_co_u32_e64 %[s1c], %[Kc], %[a1c], %[b1c], %[Kc]\n\t"
        "v_subb_co_u32_e64 %[s1d], %[Kd], %[a1d], %[b1d], %[Kd]\n\t"
        "v_add_co_u32_e64 %[t0a], %[Ma], 1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Mb], 1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Mc], 1, %[s0c]\n\t"
        "v_add_co_u32_e64 %[t0d], %[Md], 1, %[s0d]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ma], -1, %[s1a], %[Ma]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Mb], -1, %[s1b], %[Mb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Mc], -1, %[s1c], %[Mc]\n\t"
        "v_addc_co_u32_e64 %[t1d], %[Md], -1, %[s1d], %[Md]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s0d], %[s0d], %[t0d], %[Kd]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1d], %[s1d], %[t1d], %[Kd]"
        : [s0a] "=&v"(s0[0]), [s1a] "=&v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [Ma] "=&s"(M[0]), [s0b] "=&v"(s0[1]), [s1b] "=&v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]),
          [Kb] "=&s"(K[1]), [Mb] "=&s"(M[1]), [s0c] "=&v"(s0[2]), [s1c] "=&v"(s1[2]), [t0c] "=&v"(t0[2]),
          [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2]), [Mc] "=&s"(M[2]), [s0d] "=&v"(s0[3]), [s1d] "=&v"(s1[3]),
          [t0d] "=&v"(t0[3]), [t1d] "=&v"(t1[3]), [Kd] "=&s"(K[3]), [Md] "=&s"(M[3])
        : [a0a] "v"((uint32_t)a[0]), [a1a] "v"((uint32_t)(a[0] >> 32)), [b0a] "v"((uint32_t)b[0]),
          [b1a] "v"((uint32_t)(b[0] >> 32)), [a0b] "v"((uint32_t)a[1]), [a1b] "v"((uint32_t)(a[1] >> 32)),
          [b0b] "v"((uint32_t)b[1]), [b1b] "v"((uint32_t)(b[1] >> 32)), [a0c] "v"((uint32_t)a[2]),
          [a1c] "v"((uint32_t)(a[2] >> 32)), [b0c] "v"((uint32_t)b[2]), [b1c] "v"((uint32_t)(b[2] >> 32)),
          [a0d] "v"((uint32_t)a[3]), [a1d] "v"((uint32_t)(a[3] >> 32)), [b0d] "v"((uint32_t)b[3]),
          [b1d] "v"((uint32_t)(b[3] >> 32))
        : "scc");
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = gl_pack(s0[k], s1[k]);
}

template <> __device__ __forceinline__ void gl_prod_reduce_canon_n<4>(uint64_t *out, const gl_prod *x)
{
    uint64_t r[4], c[4], B[4], P[4], K[4];
    uint32_t r0[4], r1[4], t0[4], t1[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r[k]), "=&s"(c[k]) : "v"(x[k].x2), "v"(x[k].lo));
        r0[k] = (uint32_t)r[k];
        r1[k] = (uint32_t)(r[k] >> 32);
    }
    asm(
        "v_sub_co_u32_e64 %[r0a], %[Ba], %[r0a], %[x3a]\n\t"
        "v_sub_co_u32_e64 %[r0b], %[Bb], %[r0b], %[x3b]\n\t"
        "v_sub_co_u32_e64 %[r0c], %[Bc], %[r0c], %[x3c]\n\t"
        "v_sub_co_u32_e64 %[r0d], %[Bd], %[r0d], %[x3d]\n\t"
        "v_subb_co_u32_e64 %[r1a], %[Ba], %[r1a], 0, %[Ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[Bb], %[r1b], 0, %[Bb]\n\t"
        "v_subb_co_u32_e64 %[r1c], %[Bc], %[r1c], 0, %[Bc]\n\t"
        "v_subb_co_u32_e64 %[r1d], %[Bd], %[r1d], 0, %[Bd]\n\t"
        "s_andn2_b64 %[Pa], %[Ca], %[Ba]\n\t"
        "s_andn2_b64 %[Pb], %[Cb], %[Bb]\n\t"
        "s_andn2_b64 %[Pc], %[Cc], %[Bc]\n\t"
        "s_andn2_b64 %[Pd], %[Cd], %[Bd]\n\t"
        "s_andn2_b64 %[Ba], %[Ba], %[Ca]\n\t"
        "s_andn2_b64 %[Bb], %[Bb], %[Cb]\n\t"
        "s_andn2_b64 %[Bc], %[Bc], %[Cc]\n\t"
        "s_andn2_b64 %[Bd], %[Bd], %[Cd]\n\t"
        "v_cndmask_b32_e64 %[t0a], 0, -1, %[Pa]\n\t"
        "v_cndmask_b32_e64 %[t0b], 0, -1, %[Pb]\n\t"
        "v_cndmask_b32_e64 %[t0c], 0, -1, %[Pc]\n\t"
        "v_cndmask_b32_e64 %[t0d], 0, -1, %[Pd]\n\t"
        "v_cndmask_b32_e64 %[t1a], 0, -1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t1b], 0, -1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[t1c], 0, -1, %[Bc]\n\t"
        "v_cndmask_b32_e64 %[t1d], 0, -1, %[Bd]\n\t"
        "v_cndmask_b32_e64 %[t0a], %[t0a], 1, %[Ba]\n\t"
        "v_cndmask_b32_e64 %[t0b], %[t0b], 1, %[Bb]\n\t"
        "v_cndmask_b32_e64 %[t0c], %[t0c], 1, %[Bc]\n\t"
        "v_cndmask_b32_e64 %[t0d], %[t0d], 1, %[Bd]\n\t"
        "v_add_co_u32_e64 %[r0a], %[Ca], %[r0a], %[t0a]\n\t"
        "v_add_co_u32_e64 %[r0b], %[Cb], %[r0b], %[t0b]\n\t"
        "v_add_co_u32_e64 %[r0c], %[Cc], %[r0c], %[t0c]\n\t"
        "v_add_co_u32_e64 %[r0d], %[Cd], %[r0d], %[t0d]\n\t"
        "v_addc_co_u32_e64 %[r1a], %[Ca], %[r1a], %[t1a], %[Ca]\n\t"
        "v_addc_co_u32_e64 %[r1b], %[Cb], %[r1b], %[t1b], %[Cb]\n\t"
        "v_addc_co_u32_e64 %[r1c], %[Cc], %[r1c], %[t1c], %[Cc]\n\t"
        "v_addc_co_u32_e64 %[r1d], %[Cd], %[r1d], %[t1d], %[Cd]"
        : [r0a] "+v"(r0[0]), [r1a] "+v"(r1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ba] "=&s"(B[0]),
          [Pa] "=&s"(P[0]), [Ca] "+s"(c[0]), [r0b] "+v"(r0[1]), [r1b] "+v"(r1[1]), [t0b] "=&v"(t0[1]),
          [t1b] "=&v"(t1[1]), [Bb] "=&s"(B[1]), [Pb] "=&s"(P[1]), [Cb] "+s"(c[1]), [r0c] "+v"(r0[2]),
          [r1c] "+v"(r1[2]), [t0c] "=&v"(t0[2]), [t1c] "=&v"(t1[2]), [Bc] "=&s"(B[2]), [Pc] "=&s"(P[2]),
          [Cc] "+s"(c[2]), [r0d] "+v"(r0[3]), [r1d] "+v"(r1[3]), [t0d] "=&v"(t0[3]), [t1d] "=&v"(t1[3]),
          [Bd] "=&s"(B[3]), [Pd] "=&s"(P[3]), [Cd] "+s"(c[3])
        : [x3a] "v"(x[0].x3), [x3b] "v"(x[1].x3), [x3c] "v"(x[2].x3), [x3d] "v"(x[3].x3)
        : "scc");
    uint32_t *s0 = r0, *s1 = r1;
    asm(
        "v_add_co_u32_e64 %[t0a], %[Ka], -1, %[s0a]\n\t"
        "v_add_co_u32_e64 %[t0b], %[Kb], -1, %[s0b]\n\t"
        "v_add_co_u32_e64 %[t0c], %[Kc], -1, %[s0c]\n\t"
        "v_add_co_u32_e64 %[t0d], %[Kd], -1, %[s0d]\n\t"
        "v_addc_co_u32_e64 %[t1a], %[Ka], 0, %[s1a], %[Ka]\n\t"
        "v_addc_co_u32_e64 %[t1b], %[Kb], 0, %[s1b], %[Kb]\n\t"
        "v_addc_co_u32_e64 %[t1c], %[Kc], 0, %[s1c], %[Kc]\n\t"
        "v_addc_co_u32_e64 %[t1d], %[Kd], 0, %[s1d], %[Kd]\n\t"
        "v_cndmask_b32_e64 %[s0a], %[s0a], %[t0a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s0b], %[s0b], %[t0b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s0c], %[s0c], %[t0c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s0d], %[s0d], %[t0d], %[Kd]\n\t"
        "v_cndmask_b32_e64 %[s1a], %[s1a], %[t1a], %[Ka]\n\t"
        "v_cndmask_b32_e64 %[s1b], %[s1b], %[t1b], %[Kb]\n\t"
        "v_cndmask_b32_e64 %[s1c], %[s1c], %[t1c], %[Kc]\n\t"
        "v_cndmask_b32_e64 %[s1d], %[s1d], %[t1d], %[Kd]"
        : [s0a] "+v"(s0[0]), [s1a] "+v"(s1[0]), [t0a] "=&v"(t0[0]), [t1a] "=&v"(t1[0]), [Ka] "=&s"(K[0]),
          [s0b] "+v"(s0[1]), [s1b] "+v"(s1[1]), [t0b] "=&v"(t0[1]), [t1b] "=&v"(t1[1]), [Kb] "=&s"(K[1]),
          [s0c] "+v"(s0[2]), [s1c] "+v"(s1[2]), [t0c] "=&v"(t0[2]), [t1c] "=&v"(t1[2]), [Kc] "=&s"(K[2]),
          [s0d] "+v"(s0[3]), [s1d] "+v"(s1[3]), [t0d] "=&v"(t0[3]), [t1d] "=&v"(t1[3]), [Kd] "=&s"(K[3])
        :
        : "scc");
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = gl_pack(r0[k], r1[k]);
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
// out[k] = a[k] op b[k] for k < N, in interleaved groups of GL_LANES lanes
#ifndef GL_LANES
#define GL_LANES 3
#endif
template <int N>
__device__ __forceinline__ void gl_add_v(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    constexpr int G = GL_LANES;
#pragma unroll
    for (int k = 0; k + G <= N; k += G) gl_add_n<G>(out + k, a + k, b + k);
    if constexpr (N % G) gl_add_n<N % G>(out + N / G * G, a + N / G * G, b + N / G * G);
}
template <int N>
__device__ __forceinline__ void gl_sub_v(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
    constexpr int G = GL_LANES;
#pragma unroll
    for (int k = 0; k + G <= N; k += G) gl_sub_n<G>(out + k, a + k, b + k);
    if constexpr (N % G) gl_sub_n<N % G>(out + N / G * G, a + N / G * G, b + N / G * G);
}
// canonical products out[k] = a[k] * b[k] (any 64-bit inputs; out may alias a or
// b): carry-form products three lanes at a time, then one canonicalisation
template <int N>
__device__ __forceinline__ void gl_mul_v(uint64_t *out, const uint64_t *a, const uint64_t *b)
{
#pragma unroll
    for (int k = 0; k + 3 <= N; k += 3) {
        gl_prodk x[3];
        gl_prodk3(a + k, b + k, 0, 0, x);
        gl_reducek3(x, out + k);
        gl_canon3(out + k);
    }
#pragma unroll
    for (int k = N / 3 * 3; k < N; k++) out[k] = gl_canon_dev(gl_mulk1(a[k], b[k], 0, 0));
}

// ---- multiplication by a power of two ------------------------------------
// x * 2^S mod p for 0 < S < 96 and any 64-bit x, with no 64x64 product. The roots
// of unity of order <= 64 are powers of two (w_64 = 8), so inside a 16-point DFT
// every twiddle is w_16^k = 2^(12k) or w_16^-k = -2^(96 - 12k).
//   S < 64:  x 2^S = L + 2^64 H with L = x << S and H = x >> (64 - S) < 2^S;
//            2^64 = e = 2^32 - 1 and 2^96 = -1 give L + e H0 - H1.
//   S >= 64: with t = S - 64, x 2^S = (x 2^t) 2^64 and 2^128 = -2^32 give
//            e L0 - Q, L0 = (x << t) mod 2^32 and Q = x >> (32 - t) < 2^(32 + t).
// Both are r = A e + lo - (B1:B0) (A, B0, B1 32-bit words, lo a 64-bit word): one
// v_mad_u64_u32 with carry-out c, a 64-bit subtraction with borrow-out b, then
// + e for c & !b and - e for b & !c. Neither correction wraps: c = 1 leaves
// r < (2^32 - 1)^2 before the subtraction; b = 1 with c = 0 leaves r >= 2^64 - 2^32
// (S < 64, where B1 = 0) or r >= 2^64 - 2^(32 + t) (S >= 64, where c = 0).
template <int S>
__device__ __forceinline__ void gl_pow2_terms(uint64_t x, uint32_t &A, uint64_t &lo, uint32_t &B0, uint32_t &B1)
{
    static_assert(S > 0 && S < 96, "power of two out of range");
    if constexpr (S < 64) {
        uint64_t h = x >> (64 - S);
        A = (uint32_t)h;
        lo = x << S;
        B0 = (uint32_t)(h >> 32);
        B1 = 0;
    } else {
        constexpr int t = S - 64;
        uint64_t q = x >> (32 - t);
        A = (uint32_t)x << t;
        lo = 0;
        B0 = (uint32_t)q;
        B1 = (uint32_t)(q >> 32);
    }
}

// Three lanes of r = A e + lo - (B1:B0) with the correction above (congruent
// 64-bit words); every carry mask is read at least two instructions after its write.
__device__ __forceinline__ void gl_pow2_reduce3(const uint32_t *A, const uint64_t *lo, const uint32_t *B0,
                                                const uint32_t *B1, uint64_t *out)
{
    uint64_t r[3], c[3];
    asm("v_mad_u64_u32 %0, %3, %6, -1, %9\n\t"
        "v_mad_u64_u32 %1, %4, %7, -1, %10\n\t"
        "v_mad_u64_u32 %2, %5, %8, -1, %11"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&s"(c[0]), "=&s"(c[1]), "=&s"(c[2])
        : "v"(A[0]), "v"(A[1]), "v"(A[2]), "v"(lo[0]), "v"(lo[1]), "v"(lo[2]));
    uint32_t r0[3], r1[3], tm[3], tn[3];
    uint64_t b[3], m[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r0[i] = (uint32_t)r[i];
        r1[i] = (uint32_t)(r[i] >> 32);
    }
    asm("v_sub_co_u32_e64 %[r0a], %[ba], %[r0a], %[x0a]\n\t"
        "v_sub_co_u32_e64 %[r0b], %[bb], %[r0b], %[x0b]\n\t"
        "v_sub_co_u32_e64 %[r0c], %[bc], %[r0c], %[x0c]\n\t"
        "v_subb_co_u32_e64 %[r1a], %[ba], %[r1a], %[x1a], %[ba]\n\t"
        "v_subb_co_u32_e64 %[r1b], %[bb], %[r1b], %[x1b], %[bb]\n\t"
        "v_subb_co_u32_e64 %[r1c], %[bc], %[r1c], %[x1c], %[bc]\n\t"
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
        : [x0a] "v"(B0[0]), [x0b] "v"(B0[1]), [x0c] "v"(B0[2]), [x1a] "v"(B1[0]), [x1b] "v"(B1[1]),
          [x1c] "v"(B1[2])
        : "scc");
#pragma unroll
    for (int i = 0; i < 3; i++) out[i] = gl_pack(r0[i], r1[i]);
}

// One lane of gl_pow2_reduce3, hazard gaps filled with s_nop.
__device__ __forceinline__ uint64_t gl_pow2_reduce1(uint32_t A, uint64_t lo, uint32_t B0, uint32_t B1)
{
    uint64_t r, c, bm, m;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"(A), "v"(lo));
    uint32_t r0 = (uint32_t)r, r1 = (uint32_t)(r >> 32), tm, tn;
    asm("v_sub_co_u32_e64 %[r0], %[b], %[r0], %[x0]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[r1], %[b], %[r1], %[x1], %[b]\n\t"
        "s_nop 1\n\t"
        "s_andn2_b64 %[m], %[c], %[b]\n\t"
        "s_andn2_b64 %[b], %[b], %[c]\n\t"
        "v_cndmask_b32_e64 %[tm], 0, -1, %[m]\n\t"
        "v_cndmask_b32_e64 %[tn], 0, -1, %[b]\n\t"
        "v_addc_co_u32_e64 %[r0], %[c], %[r0], %[tm], %[b]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[r1], %[c], %[r1], %[tn], %[c]"
        : [r0] "+v"(r0), [r1] "+v"(r1), [tm] "=&v"(tm), [tn] "=&v"(tn), [b] "=&s"(bm), [m] "=&s"(m), [c] "+s"(c)
        : [x0] "v"(B0), [x1] "v"(B1)
        : "scc");
    return gl_pack(r0, r1);
}

// canonical out[k] = x[k] * 2^S_k for three lanes (out may alias x)
template <int S0, int S1, int S2>
__device__ __forceinline__ void gl_mul_pow2_3(const uint64_t *x, uint64_t *out)
{
    uint32_t A[3], B0[3], B1[3];
    uint64_t lo[3];
    gl_pow2_terms<S0>(x[0], A[0], lo[0], B0[0], B1[0]);
    gl_pow2_terms<S1>(x[1], A[1], lo[1], B0[1], B1[1]);
    gl_pow2_terms<S2>(x[2], A[2], lo[2], B0[2], B1[2]);
    gl_pow2_reduce3(A, lo, B0, B1, out);
    gl_canon3(out);
}

// canonical x * 2^S
template <int S>
__device__ __forceinline__ uint64_t gl_mul_pow2(uint64_t x)
{
    uint32_t A, B0, B1;
    uint64_t lo;
    gl_pow2_terms<S>(x, A, lo, B0, B1);
    return gl_canon_dev(gl_pow2_reduce1(A, lo, B0, B1));
}

#else  // host pass: declarations only (kernels are compiled for the device)
template <int N> __device__ void gl_add_v(uint64_t *out, const uint64_t *a, const uint64_t *b);
template <int N> __device__ void gl_sub_v(uint64_t *out, const uint64_t *a, const uint64_t *b);
template <int N> __device__ void gl_mul_v(uint64_t *out, const uint64_t *a, const uint64_t *b);
template <int S0, int S1, int S2> __device__ void gl_mul_pow2_3(const uint64_t *x, uint64_t *out);
template <int S> __device__ uint64_t gl_mul_pow2(uint64_t x);
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
