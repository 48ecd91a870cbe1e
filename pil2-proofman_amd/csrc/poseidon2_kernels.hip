// Poseidon2-Goldilocks Merkle commitment kernels for gfx950.
//
// Semantics:
//   permutation   pil2-stark/src/goldilocks/src/poseidon2_goldilocks.cpp:10-41
//                 (external layer, 4 full rounds, 22 partial rounds, 4 full rounds;
//                  external layer = M4 on each 4-block + column sums,
//                  poseidon2_goldilocks.hpp:186-231)
//   leaf sponge   poseidon2_goldilocks.cpp:43-77 (rate W-4, capacity 4, capacity
//                 of block k>0 = state[0..4) of block k-1, zero-padded rate,
//                 empty row -> zero digest)
//   tree          poseidon2_goldilocks.cpp:130-175 (levels padded with zero digests
//                 to a multiple of arity, parent = first 4 words of permute(children))
//   width         W = 4 * arity (merkleTreeGL.cpp:316-336)
//
// One permutation per lane: the state lives in VGPRs, round constants are
// wave-uniform and come from __constant__ memory through scalar loads.
#include "gl_field.h"
#include "p2_arith.h"
#include "p2_constants.h"
#include "poseidon2_kernels.h"

__constant__ uint64_t c_rc8[86] = P2_RC8_INIT;
__constant__ uint64_t c_rc12[118] = P2_RC12_INIT;
__constant__ uint64_t c_rc16[150] = P2_RC16_INIT;
__constant__ uint64_t c_d8[8] = P2_DIAG8_INIT;
__constant__ uint64_t c_d12[12] = P2_DIAG12_INIT;
__constant__ uint64_t c_d16[16] = P2_DIAG16_INIT;

template <int W> struct P2Tab;
template <> struct P2Tab<8> {
    static __device__ __forceinline__ uint64_t rc(int i) { return c_rc8[i]; }
    static __device__ __forceinline__ uint64_t d(int i) { return c_d8[i]; }
};
template <> struct P2Tab<12> {
    static __device__ __forceinline__ uint64_t rc(int i) { return c_rc12[i]; }
    static __device__ __forceinline__ uint64_t d(int i) { return c_d12[i]; }
};
template <> struct P2Tab<16> {
    static __device__ __forceinline__ uint64_t rc(int i) { return c_rc16[i]; }
    static __device__ __forceinline__ uint64_t d(int i) { return c_d16[i]; }
};

// External layer M_E (poseidon2_goldilocks.hpp:186-231): M4 on every 4-block,
// then every element adds the sum of its column over the blocks. It runs on the
// 32-bit halves of the input words separately, in 64-bit accumulators: the
// output is in halves form with each accumulator below 2^39 (M4 rows sum to at
// most 16, so a block output is below 2^36; at most 4 blocks plus the element
// itself make a column total below 2^39).
// M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] as additions and shifts:
// t0 = x0+x1, t1 = x2+x3, t2 = 2x1+t1, t3 = 2x3+t0, t4 = 4t1+t3, t5 = 4t0+t2,
// (x0, x1, x2, x3) <- (t3+t5, t5, t2+t4, t4).
// M4 on the 32-bit halves a[0..4) and b[0..4) of two blocks, in 64-bit sums:
// t0 and t1 are 32-bit additions with carry (no zero-extended operands), t2
// and t3 multiply-adds with a 32-bit operand, the rest 64-bit shift-adds.
__device__ __forceinline__ void m4_halves2(const uint32_t *a, const uint32_t *b, uint64_t *ya, uint64_t *yb)
{
    uint32_t l[4], h[4];
    uint64_t c[4];
    asm("v_add_co_u32_e64 %0, %8, %12, %13\n\t"
        "v_add_co_u32_e64 %1, %9, %14, %15\n\t"
        "v_add_co_u32_e64 %2, %10, %16, %17\n\t"
        "v_add_co_u32_e64 %3, %11, %18, %19\n\t"
        "v_addc_co_u32_e64 %4, %8, 0, 0, %8\n\t"
        "v_addc_co_u32_e64 %5, %9, 0, 0, %9\n\t"
        "v_addc_co_u32_e64 %6, %10, 0, 0, %10\n\t"
        "v_addc_co_u32_e64 %7, %11, 0, 0, %11"
        : "=&v"(l[0]), "=&v"(l[1]), "=&v"(l[2]), "=&v"(l[3]), "=&v"(h[0]), "=&v"(h[1]), "=&v"(h[2]), "=&v"(h[3]),
          "=&s"(c[0]), "=&s"(c[1]), "=&s"(c[2]), "=&s"(c[3])
        : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(b[0]), "v"(b[1]), "v"(b[2]), "v"(b[3]));
    uint64_t t0a = gl_pack(l[0], h[0]), t1a = gl_pack(l[1], h[1]), t0b = gl_pack(l[2], h[2]), t1b = gl_pack(l[3], h[3]);
    uint64_t t2a, t3a, t2b, t3b, d[4];
    asm("v_mad_u64_u32 %0, %4, %8, 2, %12\n\t"
        "v_mad_u64_u32 %1, %5, %9, 2, %13\n\t"
        "v_mad_u64_u32 %2, %6, %10, 2, %14\n\t"
        "v_mad_u64_u32 %3, %7, %11, 2, %15"
        : "=&v"(t2a), "=&v"(t3a), "=&v"(t2b), "=&v"(t3b), "=&s"(d[0]), "=&s"(d[1]), "=&s"(d[2]), "=&s"(d[3])
        : "v"(a[1]), "v"(a[3]), "v"(b[1]), "v"(b[3]), "v"(t1a), "v"(t0a), "v"(t1b), "v"(t0b));
    uint64_t t4a = (t1a << 2) + t3a, t5a = (t0a << 2) + t2a, t4b = (t1b << 2) + t3b, t5b = (t0b << 2) + t2b;
    ya[0] = t3a + t5a;
    ya[1] = t5a;
    ya[2] = t2a + t4a;
    ya[3] = t4a;
    yb[0] = t3b + t5b;
    yb[1] = t5b;
    yb[2] = t2b + t4b;
    yb[3] = t4b;
}

template <int W>
__device__ __forceinline__ void ext_layer_hz(const uint64_t *s, hz *t)
{
    constexpr int NB = W / 4;
    uint64_t y[2][W], col[2][4];
    uint32_t hv[2][W];
#pragma unroll
    for (int i = 0; i < W; i++) {
        hv[0][i] = (uint32_t)s[i];
        hv[1][i] = (uint32_t)(s[i] >> 32);
    }
    // the 2*NB half-blocks, two at a time
#pragma unroll
    for (int k = 0; k < NB; k++) m4_halves2(&hv[0][4 * k], &hv[1][4 * k], &y[0][4 * k], &y[1][4 * k]);
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            col[h][k] = y[h][k];
#pragma unroll
            for (int j = 1; j < NB; j++) col[h][k] += y[h][4 * j + k];
        }
    }
#pragma unroll
    for (int i = 0; i < W; i++) {
        t[i].lo = y[0][i] + col[0][i & 3];
        t[i].hi = y[1][i] + col[1][i & 3];
    }
}

// x^7 (any 64-bit word in, congruent word out)
__device__ __forceinline__ uint64_t pow7(uint64_t x)
{
    uint64_t x2 = p2_mul(x, x);
    uint64_t x3 = p2_mul(x2, x);
    uint64_t x4 = p2_mul(x2, x2);
    return p2_mul(x3, x4);
}

// x^7 on three lanes, each product interleaved across the lanes
__device__ __forceinline__ void pow7_3(const uint64_t *x, uint64_t *out)
{
    uint64_t x2[3], x3[3], x4[3];
    p2_mul3(x, x, x2);
    p2_mul3(x2, x, x3);
    p2_mul3(x2, x2, x4);
    p2_mul3(x3, x4, out);
}

// s[i] = fold(t[i]) for all lanes, three at a time
template <int W>
__device__ __forceinline__ void fold_all(const hz *t, uint64_t *s)
{
    constexpr int G = W / 3 * 3;
#pragma unroll
    for (int i = 0; i < G; i += 3) hz_fold3(t + i, s + i);
#pragma unroll
    for (int i = G; i < W; i++) s[i] = hz_fold(t[i]);
}

// Full round on a halves-form state t: S-box(t + rc) on every element, then M_E.
// The round constant's halves go into the accumulators (still below 2^40).
template <int W>
__device__ __forceinline__ void full_round(hz *t, int rc_base)
{
    constexpr int G = W / 3 * 3;  // lanes handled three at a time
    uint64_t x[W], s[W];
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t rc = P2Tab<W>::rc(rc_base + i);
        t[i].lo += (uint32_t)rc;
        t[i].hi += rc >> 32;
    }
    fold_all<W>(t, x);
#pragma unroll
    for (int i = 0; i < G; i += 3) pow7_3(x + i, s + i);
#pragma unroll
    for (int i = G; i < W; i++) s[i] = pow7(x[i]);
    ext_layer_hz<W>(s, t);
}

// Sum of W words in halves form: the low and the high 32-bit halves are summed
// separately, each as one 32-bit addition with carry and then a chain of
// v_mad_u64_u32 (half * 1 + accumulator), which takes the 32-bit half as it is
// (no zero-extended copy). Each accumulator stays below W * 2^32.
template <int W>
__device__ __forceinline__ hz sum_halves(const uint64_t *s)
{
    hz a;
    a.lo = (uint64_t)(uint32_t)s[0] + (uint32_t)s[1];
    a.hi = (s[0] >> 32) + (s[1] >> 32);
#pragma unroll
    for (int i = 2; i < W; i++) {
        uint64_t nl, nh, d0, d1;
        asm("v_mad_u64_u32 %0, %2, %4, 1, %6\n\t"
            "v_mad_u64_u32 %1, %3, %5, 1, %7"
            : "=&v"(nl), "=&v"(nh), "=&s"(d0), "=&s"(d1)
            : "v"((uint32_t)s[i]), "v"((uint32_t)(s[i] >> 32)), "v"(a.lo), "v"(a.hi));
        a.lo = nl;
        a.hi = nh;
    }
    return a;
}

template <int W>
__device__ __forceinline__ void p2_permute(uint64_t *s)
{
    // Round loops stay rolled: the unrolled permutation would not fit the
    // instruction cache, and per-round constants are wave-uniform scalar loads.
    constexpr int G = W / 3 * 3;
    hz t[W];
    ext_layer_hz<W>(s, t);
#pragma unroll 1
    for (int r = 0; r < 4; r++) full_round<W>(t, r * W);
    fold_all<W>(t, s);
    // partial rounds: S-box on element 0, then s <- diag(d) s + sum(s)
#pragma unroll 1
    for (int r = 0; r < 22; r++) {
        lz rc0 = lz_of(P2Tab<W>::rc(4 * W + r)), u, x0 = lz_of(s[0]);
        lz_add_n<1>(&u, &x0, &rc0);
        s[0] = pow7(lz_fold(u));
        uint64_t sum = hz_fold(sum_halves<W>(s));
        uint32_t s_lo = (uint32_t)sum, s_hi = (uint32_t)(sum >> 32);
        uint64_t d[W];
#pragma unroll
        for (int i = 0; i < W; i++) d[i] = P2Tab<W>::d(i);
#pragma unroll
        for (int i = 0; i < G; i += 3) p2_mul_add3(s + i, d + i, s_lo, s_hi, s + i);
#pragma unroll
        for (int i = G; i < W; i++) s[i] = p2_mul_add_u(s[i], d[i], s_lo, s_hi);
    }
#pragma unroll
    for (int i = 0; i < W; i++) t[i] = {(uint32_t)s[i], s[i] >> 32};
#pragma unroll 1
    for (int r = 0; r < 4; r++) full_round<W>(t, 4 * W + 22 + r * W);
    fold_all<W>(t, s);
}

// ------------------------------------------------------------ leaves ----
// Row r of the source: element c at src[r*rs + c*cs]. Digest to out[r*4 .. r*4+3].
// carry != 0 continues a sponge over later columns: the capacity starts from
// the digest words already in out (the state words 0..3 after the previous
// columns, which is what the next block's capacity takes), so absorbing a row's
// columns in groups whose widths are multiples of the rate gives the one-call
// digest.
template <int W>
__global__ __launch_bounds__(P2_THREADS) void k_p2_leaves(const uint64_t *__restrict__ src, uint64_t rs, uint64_t cs,
                                                          uint64_t ncols, uint64_t nrows, uint64_t *__restrict__ out,
                                                          uint32_t carry)
{
    constexpr int RATE = W - 4;
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    uint64_t s[W];
    const uint64_t *row = src + r * rs;
#pragma unroll
    for (int i = RATE; i < W; i++) s[i] = carry ? out[r * 4 + (i - RATE)] : 0;
    if (ncols == 0) {
        if (!carry) {
#pragma unroll
            for (int i = 0; i < 4; i++) out[r * 4 + i] = 0;
        }
        return;
    }
    for (uint64_t c0 = 0; c0 < ncols; c0 += RATE) {
        if (c0 > 0) {
#pragma unroll
            for (int i = 0; i < 4; i++) s[RATE + i] = s[i];
        }
        if (c0 + RATE <= ncols) {
#pragma unroll
            for (int i = 0; i < RATE; i++) s[i] = row[(c0 + i) * cs];
        } else {
#pragma unroll
            for (int i = 0; i < RATE; i++) s[i] = (c0 + i < ncols) ? row[(c0 + i) * cs] : 0;
        }
        p2_permute<W>(s);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[r * 4 + i] = gl_canon(s[i]);
}

// ------------------------------------------------------------- nodes ----
// One level: children = `pending` digests at in (pad digests beyond it are
// zero), parents written to out (= in + (pending + pad)*4). The same kernel
// also writes the zero pad digests after the children.
template <int W>
__global__ __launch_bounds__(P2_THREADS) void k_p2_nodes(uint64_t *__restrict__ tree, uint64_t in_off, uint64_t pending,
                                                         uint64_t pad, uint64_t nnext)
{
    constexpr int ARITY = W / 4;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnext) return;
    uint64_t s[W];
    const uint64_t *ch = tree + in_off + i * W;
#pragma unroll
    for (int a = 0; a < ARITY; a++) {
        bool real = i * ARITY + a < pending;
#pragma unroll
        for (int k = 0; k < 4; k++) s[a * 4 + k] = real ? ch[a * 4 + k] : 0;
    }
    if (i == nnext - 1) {
        for (uint64_t z = 0; z < pad * 4; z++) tree[in_off + pending * 4 + z] = 0;
    }
    p2_permute<W>(s);
    uint64_t *o = tree + in_off + (pending + pad) * 4 + i * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = gl_canon(s[k]);
}

// ------------------------------------------------- cooperative levels ----
// Upper tree levels have too few parents to fill the GPU with one permutation
// per lane, so their time is the latency of one permutation. Here the 16 lanes
// of a DPP row share one permutation: lane i holds state word i and lanes >= W
// hold zero. Full rounds run all S-boxes at once; M4 reads the four words of
// the lane's quad through DPP quad broadcasts; column sums and the state sum of
// the partial rounds are DPP row rotations.

// DPP reads of a VGPR need two wait states after its VALU write; values made by
// inline asm pass through this barrier-nop first (the "+v" ties the order).
__device__ __forceinline__ void dpp_ready(uint32_t &a, uint32_t &b, uint32_t &c)
{
    asm volatile("s_nop 1" : "+v"(a), "+v"(b), "+v"(c));
}

// the same barrier for two 64-bit values
__device__ __forceinline__ void dpp_ready64(uint64_t &a, uint64_t &b)
{
    asm volatile("s_nop 1" : "+v"(a), "+v"(b));
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return gl_pack(lo, hi);
}

constexpr int dpp_quad_bcast(int k) { return k | (k << 2) | (k << 4) | (k << 6); }
constexpr int dpp_row_ror(int n) { return 0x120 + n; }

__device__ __forceinline__ lz lz_add1(lz a, lz b)
{
    lz r;
    lz_add_n<1>(&r, &a, &b);
    return r;
}

// M_E on the row in halves form: the quad's four words are broadcast as 32-bit
// halves, M4 runs on each half in 64-bit sums, and the column sum over the
// blocks is two DPP row rotations of the 64-bit accumulators. Lane i gets
// (M4 x_quad)[i mod 4] + its column sum; lanes >= W get zero.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

template <int W>
__device__ __forceinline__ hz ext_coop_hz(uint64_t x, uint32_t i, bool live)
{
    uint64_t xv = x, dummy = 0;
    dpp_ready64(xv, dummy);
    uint32_t xl = (uint32_t)xv, xh = (uint32_t)(xv >> 32);
    uint32_t q = i & 3;
    uint64_t y[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t v = h ? xh : xl;
        uint64_t a0 = dpp_u32<dpp_quad_bcast(0)>(v), a1 = dpp_u32<dpp_quad_bcast(1)>(v);
        uint64_t a2 = dpp_u32<dpp_quad_bcast(2)>(v), a3 = dpp_u32<dpp_quad_bcast(3)>(v);
        uint64_t t0 = a0 + a1, t1 = a2 + a3;
        uint64_t t2 = (a1 << 1) + t1, t3 = (a3 << 1) + t0;
        uint64_t t4 = (t1 << 2) + t3, t5 = (t0 << 2) + t2;
        y[h] = q == 0 ? t3 + t5 : q == 1 ? t5 : q == 2 ? t2 + t4 : t4;
    }
    hz c = {y[0], y[1]};
    dpp_ready64(c.lo, c.hi);
    c.lo += dpp_u64<dpp_row_ror(4)>(c.lo);
    c.hi += dpp_u64<dpp_row_ror(4)>(c.hi);
    dpp_ready64(c.lo, c.hi);
    c.lo += dpp_u64<dpp_row_ror(8)>(c.lo);
    c.hi += dpp_u64<dpp_row_ror(8)>(c.hi);
    return live ? hz{y[0] + c.lo, y[1] + c.hi} : hz{0, 0};
}

template <int W>
__device__ __forceinline__ uint64_t p2_permute_coop(uint64_t s, uint32_t i)
{
    const bool live = i < (uint32_t)W;
    const uint32_t li = live ? i : 0;
    hz t = ext_coop_hz<W>(s, i, live);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
        uint64_t rc = live ? P2Tab<W>::rc(r * W + li) : 0;
        t.lo += (uint32_t)rc;
        t.hi += rc >> 32;
        t = ext_coop_hz<W>(pow7(hz_fold(t)), i, live);
    }
    s = hz_fold(t);
    const uint64_t d = live ? P2Tab<W>::d(li) : 0;
#pragma unroll 1
    for (int r = 0; r < 22; r++) {
        uint64_t x0 = pow7(lz_fold(lz_add1(lz_of(s), lz_of(P2Tab<W>::rc(4 * W + r)))));
        s = i == 0 ? x0 : s;
        // row sum in halves form: the two 32-bit halves in 64-bit accumulators
        // (below 16 * 2^32 after the four rotations), one 64-bit add per step
        hz v = {(uint32_t)s, s >> 32};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dpp_ready64(v.lo, v.hi);
            switch (k) {
            case 0: v.lo += dpp_u64<dpp_row_ror(1)>(v.lo); v.hi += dpp_u64<dpp_row_ror(1)>(v.hi); break;
            case 1: v.lo += dpp_u64<dpp_row_ror(2)>(v.lo); v.hi += dpp_u64<dpp_row_ror(2)>(v.hi); break;
            case 2: v.lo += dpp_u64<dpp_row_ror(4)>(v.lo); v.hi += dpp_u64<dpp_row_ror(4)>(v.hi); break;
            default: v.lo += dpp_u64<dpp_row_ror(8)>(v.lo); v.hi += dpp_u64<dpp_row_ror(8)>(v.hi); break;
            }
        }
        uint64_t sum = hz_fold(v);
        s = live ? p2_mul_add(s, d, (uint32_t)sum, (uint32_t)(sum >> 32)) : 0;
    }
    t = {(uint32_t)s, s >> 32};
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
        uint64_t rc = live ? P2Tab<W>::rc(4 * W + 22 + r * W + li) : 0;
        t.lo += (uint32_t)rc;
        t.hi += rc >> 32;
        t = ext_coop_hz<W>(pow7(hz_fold(t)), i, live);
    }
    return hz_fold(t);
}

// One tree level, 16 lanes per parent (same layout and pad rule as k_p2_nodes).
template <int W>
__global__ __launch_bounds__(P2_THREADS) void k_p2_nodes_coop(uint64_t *__restrict__ tree, uint64_t in_off,
                                                              uint64_t pending, uint64_t pad, uint64_t nnext)
{
    constexpr int ARITY = W / 4;
    uint64_t par = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    uint32_t i = threadIdx.x & 15;
    bool active = par < nnext;  // whole rows are active or not
    uint64_t s = 0;
    if (active && i < (uint32_t)W && par * ARITY + (i >> 2) < pending) s = tree[in_off + par * W + i];
    if (active && par == nnext - 1 && i == 0) {
        for (uint64_t z = 0; z < pad * 4; z++) tree[in_off + pending * 4 + z] = 0;
    }
    s = p2_permute_coop<W>(s, i);
    if (active && i < 4) tree[in_off + (pending + pad) * 4 + par * 4 + i] = gl_canon(s);
}

// Leaves of short trees, 16 lanes per row (same sponge as k_p2_leaves): lane i < RATE
// absorbs column c0 + i, capacity lanes RATE..RATE+3 take the previous output words
// 0..3 (a DPP row rotation by RATE), lanes >= W stay zero.
template <int W>
__global__ __launch_bounds__(P2_THREADS) void k_p2_leaves_coop(const uint64_t *__restrict__ src, uint64_t rs,
                                                               uint64_t cs, uint64_t ncols, uint64_t nrows,
                                                               uint64_t *__restrict__ out, uint32_t carry)
{
    constexpr int RATE = W - 4;
    uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    uint32_t i = threadIdx.x & 15;
    bool active = r < nrows;  // whole rows are active or not
    if (ncols == 0) {
        if (active && i < 4 && !carry) out[r * 4 + i] = 0;
        return;
    }
    const uint64_t *row = src + (active ? r : 0) * rs;
    uint64_t s = 0;
    for (uint64_t c0 = 0; c0 < ncols; c0 += RATE) {
        uint64_t carried = 0;
        if (c0 == 0 && carry && active && i >= (uint32_t)RATE && i < (uint32_t)W) {
            carried = out[r * 4 + (i - RATE)];  // continued sponge (see k_p2_leaves)
        } else if (c0 > 0) {  // capacity <- previous output words 0..3
            uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32), z = 0;
            dpp_ready(lo, hi, z);
            lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)lo, dpp_row_ror(RATE), 0xF, 0xF, false);
            hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)hi, dpp_row_ror(RATE), 0xF, 0xF, false);
            carried = ((uint64_t)hi << 32) | lo;
        }
        uint64_t x = 0;
        if (i < (uint32_t)RATE) {
            if (active && c0 + i < ncols) x = row[(c0 + i) * cs];
        } else if (i < (uint32_t)W) {
            x = carried;
        }
        s = p2_permute_coop<W>(x, i);
    }
    if (active && i < 4) out[r * 4 + i] = gl_canon(s);
}

// ----------------------------------------------------------- grinding ----
// Poseidon2Goldilocks<8>::grinding (poseidon2_goldilocks.cpp:177-232): nonce i
// is valid when permute([c0, c1, c2, i, 0, 0, 0, 0])[0] < 2^(64 - n_bits). One
// lane per nonce of the window [base, base + count); valid nonces go through an
// atomic minimum, so a window reports its smallest valid nonce.
__global__ __launch_bounds__(P2_THREADS) void k_p2_grind(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t level,
                                                         uint64_t base, uint64_t count,
                                                         unsigned long long *__restrict__ found)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    uint64_t s[8] = {c0, c1, c2, base + k, 0, 0, 0, 0};
    p2_permute<8>(s);
    if (gl_canon(s[0]) < level) atomicMin(found, (unsigned long long)(base + k));
}

// --------------------------------------------------------- transcript ----
// Fiat-Shamir transcript (TranscriptGL, transcriptGL.cpp:4-119): a sponge of
// width W = 4*arity, rate W-4, capacity 4. Inputs collect in `pending`; a full
// rate (or a read with inputs pending) permutes [pending, zero-padded |
// state[0..4)], and the whole output becomes the new state and the output
// buffer that field reads consume from the front. One lane runs it: the work is
// a few permutations per proof, so latency is all that matters.
template <int W>
__device__ void tr_update(P2TranscriptState *t)
{
    constexpr int RATE = W - 4;
    uint64_t s[W];
#pragma unroll
    for (int i = 0; i < RATE; i++) s[i] = (uint32_t)i < t->pending_cursor ? t->pending[i] : 0;
#pragma unroll
    for (int i = 0; i < 4; i++) s[RATE + i] = t->state[i];
    p2_permute<W>(s);
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t v = gl_canon(s[i]);
        t->out[i] = v;
        t->state[i] = v;
    }
#pragma unroll
    for (int i = 0; i < RATE; i++) t->pending[i] = 0;
    t->pending_cursor = 0;
    t->out_cursor = W;
}

template <int W>
__device__ uint64_t tr_field1(P2TranscriptState *t)
{
    if (t->out_cursor == 0) tr_update<W>(t);
    uint64_t r = t->out[(W - t->out_cursor) % W];
    t->out_cursor--;
    return r;
}

// op 0: absorb n words from io; op 1: squeeze n field words into io;
// op 2: the first n state words into io (after absorbing pending inputs)
template <int W>
__global__ void k_transcript(P2TranscriptState *t, uint64_t *io, uint64_t n, int op)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (op == 0) {
        for (uint64_t k = 0; k < n; k++) {
            t->pending[t->pending_cursor++] = io[k];
            t->out_cursor = 0;
            if (t->pending_cursor == W - 4) tr_update<W>(t);
        }
    } else if (op == 1) {
        for (uint64_t k = 0; k < n; k++) io[k] = tr_field1<W>(t);
    } else {
        if (t->pending_cursor > 0) tr_update<W>(t);
        for (uint64_t k = 0; k < n; k++) io[k] = t->state[k];
    }
}

// Single permutation over `count` states stored consecutively (W words each).
template <int W>
__global__ void k_p2_permute_many(uint64_t *st, uint64_t count)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t s[W];
#pragma unroll
    for (int k = 0; k < W; k++) s[k] = st[i * W + k];
    p2_permute<W>(s);
#pragma unroll
    for (int k = 0; k < W; k++) st[i * W + k] = gl_canon(s[k]);
}

// ---------------------------------------------------- product peak ----
// Throughput calibration of the permutation's reduced product: lane i holds
// MP_CHAINS independent chains a[k] <- a[k] * b[k] (the same p2_mul3 the
// permutation runs, in groups of three) for `iters` steps; out[C*i + k] =
// a[k] * b[k]^iters (canonical), C = MP_CHAINS. Inputs: a[k] = in[2C*i + k],
// b[k] = in[2C*i + C + k].
#define MP_GROUPS 6
#define MP_CHAINS (3 * MP_GROUPS)
__global__ __launch_bounds__(256) void k_mul_peak(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                  uint64_t lanes, uint32_t iters)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lanes) return;
    uint64_t a[MP_CHAINS], b[MP_CHAINS];
#pragma unroll
    for (int k = 0; k < MP_CHAINS; k++) {
        a[k] = in[i * 2 * MP_CHAINS + k];
        b[k] = in[i * 2 * MP_CHAINS + MP_CHAINS + k];
    }
#pragma unroll 1
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int g = 0; g < MP_GROUPS; g++) p2_mul3(a + 3 * g, b + 3 * g, a + 3 * g);
    }
#pragma unroll
    for (int k = 0; k < MP_CHAINS; k++) out[i * MP_CHAINS + k] = gl_canon(a[k]);
}

// ------------------------------------------------------------ openings ----
// Group proofs (MerkleTreeGL::getGroupProof, merkleTreeGL.cpp:145-175) for many
// queries: block q writes proof q at out + q*pw: the nCols values of source row
// idx (element (r, c) at src[r*rs + c*cs]), then for every
// level above the last-level-verification cut the arity-1 sibling digests in
// ascending child order. A query index >= nrows yields an all-zero proof.
__global__ __launch_bounds__(256) void k_group_proofs(const uint64_t *__restrict__ src, uint64_t rs, uint64_t cs,
                                                      uint64_t ncols,
                                                      const uint64_t *__restrict__ tree, uint64_t nrows,
                                                      uint64_t arity, uint64_t cut, const uint64_t *__restrict__ idx,
                                                      uint64_t pw, uint64_t *__restrict__ out)
{
    uint64_t q = blockIdx.x;
    uint64_t r = idx[q];
    uint64_t *o = out + q * pw;
    if (r >= nrows) {
        for (uint64_t k = threadIdx.x; k < pw; k += blockDim.x) o[k] = 0;
        return;
    }
    for (uint64_t c = threadIdx.x; c < ncols; c += blockDim.x) o[c] = gl_canon(src[r * rs + c * cs]);
    // siblings: level by level; thread t copies word t of each level's (arity-1)*4 words
    uint64_t per = (arity - 1) * 4;
    uint64_t t = threadIdx.x;
    if (t >= per) return;
    uint64_t n = nrows, off = 0, i = r, pos = ncols;
    while (n > cut) {
        uint64_t cur = i % arity, si = i - cur;
        uint64_t sib = t / 4, w = t % 4;
        uint64_t child = sib < cur ? sib : sib + 1;  // skip the node on the path
        o[pos + t] = tree[(off + si + child) * 4 + w];
        uint64_t nn = (n + arity - 1) / arity;
        off += nn * arity;
        i /= arity;
        n = nn;
        pos += per;
    }
}

// Copies one 4-word root (used to gather the roots of a batched commit).
__global__ void k_copy_root(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst)
{
    if (threadIdx.x < 4) dst[threadIdx.x] = src[threadIdx.x];
}

// ------------------------------------------------------------ launchers ----
extern "C" int p2_launch_group_proofs(const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols, const uint64_t *tree,
                                      uint64_t nrows, uint64_t arity, uint64_t cut, const uint64_t *idx,
                                      uint64_t nq, uint64_t pw, uint64_t *out, hipStream_t st)
{
    if (nq == 0) return 0;
    hipLaunchKernelGGL(k_group_proofs, dim3((unsigned)nq), dim3(256), 0, st, src, rs, cs, ncols, tree, nrows, arity, cut,
                       idx, pw, out);
    return (int)hipGetLastError();
}

extern "C" int p2_launch_grind(const uint64_t *ch, uint64_t level, uint64_t base, uint64_t count,
                               uint64_t *found, hipStream_t st)
{
    if (count == 0) return 0;
    dim3 g((unsigned)((count + P2_THREADS - 1) / P2_THREADS));
    hipLaunchKernelGGL(k_p2_grind, g, dim3(P2_THREADS), 0, st, ch[0], ch[1], ch[2], level, base, count,
                       (unsigned long long *)found);
    return (int)hipGetLastError();
}

extern "C" int p2_launch_transcript(uint32_t width, P2TranscriptState *t, uint64_t *io, uint64_t n, int op,
                                    hipStream_t st)
{
    switch (width) {
    case 8: hipLaunchKernelGGL(k_transcript<8>, dim3(1), dim3(64), 0, st, t, io, n, op); break;
    case 12: hipLaunchKernelGGL(k_transcript<12>, dim3(1), dim3(64), 0, st, t, io, n, op); break;
    case 16: hipLaunchKernelGGL(k_transcript<16>, dim3(1), dim3(64), 0, st, t, io, n, op); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}

extern "C" int p2_launch_copy_root(const uint64_t *src, uint64_t *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_root, dim3(1), dim3(64), 0, st, src, dst);
    return (int)hipGetLastError();
}

extern "C" int p2_launch_leaves(uint32_t width, const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols,
                                uint64_t nrows, uint64_t *out, hipStream_t st)
{
    return p2_launch_leaves_carry(width, src, rs, cs, ncols, nrows, out, 0, st);
}

extern "C" int p2_launch_leaves_carry(uint32_t width, const uint64_t *src, uint64_t rs, uint64_t cs, uint64_t ncols,
                                      uint64_t nrows, uint64_t *out, uint32_t carry, hipStream_t st)
{
    if (nrows == 0) return 0;
    if (nrows <= P2_COOP_LEAF_MAX_ROWS) {  // too few rows to fill the GPU: 16 lanes per row
        dim3 g((unsigned)((nrows * 16 + P2_THREADS - 1) / P2_THREADS));
        switch (width) {
        case 8: hipLaunchKernelGGL(k_p2_leaves_coop<8>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
        case 12: hipLaunchKernelGGL(k_p2_leaves_coop<12>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
        case 16: hipLaunchKernelGGL(k_p2_leaves_coop<16>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
        default: return -1;
        }
        return (int)hipGetLastError();
    }
    dim3 g((unsigned)((nrows + P2_THREADS - 1) / P2_THREADS));
    switch (width) {
    case 8: hipLaunchKernelGGL(k_p2_leaves<8>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
    case 12: hipLaunchKernelGGL(k_p2_leaves<12>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
    case 16: hipLaunchKernelGGL(k_p2_leaves<16>, g, dim3(P2_THREADS), 0, st, src, rs, cs, ncols, nrows, out, carry); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}

extern "C" int p2_launch_tree_levels(uint32_t width, uint64_t *tree, uint64_t nleaves, uint64_t coop_max, hipStream_t st)
{
    uint64_t arity = width / 4;
    uint64_t pending = nleaves, off = 0;
    while (pending > 1) {
        uint64_t pad = (arity - pending % arity) % arity;
        uint64_t next = (pending + arity - 1) / arity;
        if (next <= coop_max) {  // latency-bound level: 16 lanes per parent
            dim3 g((unsigned)((next * 16 + P2_THREADS - 1) / P2_THREADS));
            switch (width) {
            case 8: hipLaunchKernelGGL(k_p2_nodes_coop<8>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            case 12: hipLaunchKernelGGL(k_p2_nodes_coop<12>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            case 16: hipLaunchKernelGGL(k_p2_nodes_coop<16>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            default: return -1;
            }
        } else {
            dim3 g((unsigned)((next + P2_THREADS - 1) / P2_THREADS));
            switch (width) {
            case 8: hipLaunchKernelGGL(k_p2_nodes<8>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            case 12: hipLaunchKernelGGL(k_p2_nodes<12>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            case 16: hipLaunchKernelGGL(k_p2_nodes<16>, g, dim3(P2_THREADS), 0, st, tree, off, pending, pad, next); break;
            default: return -1;
            }
        }
        off += (pending + pad) * 4;
        pending = next;
    }
    return (int)hipGetLastError();
}

extern "C" int p2_launch_mul_peak(const uint64_t *in, uint64_t *out, uint64_t lanes, uint32_t iters, hipStream_t st)
{
    if (lanes == 0) return 0;
    hipLaunchKernelGGL(k_mul_peak, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, in, out, lanes, iters);
    return (int)hipGetLastError();
}

extern "C" int p2_launch_permute(uint32_t width, uint64_t *states, uint64_t count, hipStream_t st)
{
    if (count == 0) return 0;
    dim3 g((unsigned)((count + 127) / 128));
    switch (width) {
    case 8: hipLaunchKernelGGL(k_p2_permute_many<8>, g, dim3(128), 0, st, states, count); break;
    case 12: hipLaunchKernelGGL(k_p2_permute_many<12>, g, dim3(128), 0, st, states, count); break;
    case 16: hipLaunchKernelGGL(k_p2_permute_many<16>, g, dim3(128), 0, st, states, count); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}
