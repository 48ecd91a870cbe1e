// FRI fold and FRI-layer regrouping kernels for gfx950.
//
// k_fri_fold restates FRI::fold (pil2-stark/src/starkpil/fri/fri.hpp:32-99) with
// polMulAxi (:178-186) and evalPol (:189-206): for each folded index g in
// [0, 2^cur), gather the nX = 2^(prev-cur) cubic values pol[i*2^cur + g],
// interpolate them (inverse DFT of size nX), scale coefficient i by sinv_g^i with
// sinv_g = (7^(2^(nBitsExt-prev)) * w_{2^prev}^g)^-1, and evaluate at the
// challenge with Horner. One lane per g; the nX-point DFT runs in registers.
//
// k_fri_transpose restates getTransposed (fri.hpp:209-226): aux[i*h + j] =
// pol[j*w + i] for w = 2^nextBits, h = 2^cur / w (elements of 3 words).
#include "gl_field.h"
#include "fri_kernels.h"
#include "ntt_kernels.h"

// w^-e for the roots of one split table pair (lo[e mod 2^bits] * hi[e >> bits]);
// e is a coset index already scaled to the tables' domain. Without tables
// (lo == nullptr), w_inv^e0 by squaring and multiplying.
struct RootTab {
    const uint64_t *lo, *hi;
    uint32_t bits, shift;  // e = e0 << shift
};
__device__ __forceinline__ uint64_t root_pow(const RootTab &t, uint64_t w_inv, uint64_t e0)
{
    if (t.lo) {
        uint64_t e = e0 << t.shift;
        return gl_mul(t.lo[e & ((1ull << t.bits) - 1)], t.hi[e >> t.bits]);
    }
    return gl_pow(w_inv, e0);
}

// Fold of one group from its NX = 2^LOGX cubic values a[k][i] (component k of
// the i-th value pol[i*2^cur + g], canonical): inverse DFT of size NX, scale
// coefficient i by nx_inv * sinv^i, Horner at x. Returns the folded value.
template <int LOGX>
__device__ __forceinline__ gl3 fold_group(uint64_t (&a)[3][1 << LOGX], uint64_t sinv, gl3 x, uint64_t wx_inv,
                                          uint64_t nx_inv)
{
    constexpr int NX = 1 << LOGX;
    // inverse DFT (radix-2 DIT over the bit-reversed placement, roots of w_NX^-1)
#pragma unroll
    for (int t = 0; t < LOGX; t++) {
        int h = 1 << t;
        uint64_t wstep = wx_inv;
#pragma unroll
        for (int q = t + 1; q < LOGX; q++) wstep = gl_mul(wstep, wstep);  // w_{2h}^-1
        uint64_t w = 1;
#pragma unroll
        for (int j = 0; j < h; j++) {
#pragma unroll
            for (int b = 0; b < NX; b += 2 * h) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    uint64_t u = a[k][b + j];
                    uint64_t v = j ? gl_mul(a[k][b + j + h], w) : a[k][b + j + h];
                    a[k][b + j] = gl_add(u, v);
                    a[k][b + j + h] = gl_sub(u, v);
                }
            }
            w = gl_mul(w, wstep);
        }
    }
    // coefficient i scaled by nx_inv * sinv^i (polMulAxi), then evalPol at x
    uint64_t r = nx_inv;
#pragma unroll
    for (int i = 0; i < NX; i++) {
#pragma unroll
        for (int k = 0; k < 3; k++) a[k][i] = gl_mul(a[k][i], r);
        r = gl_mul(r, sinv);
    }
    gl3 acc = {{a[0][NX - 1], a[1][NX - 1], a[2][NX - 1]}};
#pragma unroll
    for (int i = NX - 2; i >= 0; i--) {
        gl3 c = {{a[0][i], a[1][i], a[2][i]}};
        acc = gl3_add(gl3_mul(acc, x), c);
    }
    return acc;
}

// Loads group g's values pol[i*n2 + g] (i < 2^LOGX) in bit-reversed order, so the
// DIT of fold_group yields natural order.
template <int LOGX>
__device__ __forceinline__ void load_group(uint64_t (&a)[3][1 << LOGX], const uint64_t *__restrict__ pol, uint64_t n2,
                                           uint64_t g)
{
#pragma unroll
    for (int i = 0; i < (1 << LOGX); i++) {
        int p = (int)bitrev32((uint32_t)i, LOGX);
#pragma unroll
        for (int k = 0; k < 3; k++) a[k][p] = gl_canon(pol[((uint64_t)i * n2 + g) * 3 + k]);
    }
}

__device__ __forceinline__ gl3 load_challenge(uint64_t ch0, uint64_t ch1, uint64_t ch2, const uint64_t *chp)
{
    if (chp) return {{gl_canon(chp[0]), gl_canon(chp[1]), gl_canon(chp[2])}};
    return {{ch0, ch1, ch2}};
}

// The challenge is (ch0, ch1, ch2), or the three words at chp when chp is not
// null (device memory read at run time, so a captured graph folds with the
// buffer's current contents).
template <int LOGX>
__global__ __launch_bounds__(256) void k_fri_fold(uint64_t *__restrict__ out, const uint64_t *__restrict__ pol,
                                                  uint32_t cur_bits, uint64_t shift_inv, uint64_t w_inv,
                                                  uint64_t ch0, uint64_t ch1, uint64_t ch2,
                                                  const uint64_t *__restrict__ chp, uint64_t wx_inv,
                                                  uint64_t nx_inv, RootTab rt)
{
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t n2 = 1ull << cur_bits;
    if (g >= n2) return;
    gl3 x = load_challenge(ch0, ch1, ch2, chp);
    uint64_t a[3][1 << LOGX];
    load_group<LOGX>(a, pol, n2, g);
    // sinv_g = shift_inv * w_inv^g
    gl3 r = fold_group<LOGX>(a, gl_mul(shift_inv, root_pow(rt, w_inv, g)), x, wx_inv, nx_inv);
#pragma unroll
    for (int k = 0; k < 3; k++) out[g * 3 + k] = r.v[k];
}

// Fold ratios 2^6 .. 2^10 as two folds in one kernel. With x^nX = (x^32)^(nX/32),
// the FRI fold of ratio nX at challenge c equals a fold of ratio 32 at c
// (prev -> prev-5 bits, on the coset shift s and w_{2^prev}) followed by a fold of
// ratio nB = nX/32 at c^32 (prev-5 -> cur bits, on the coset s^32 and
// w_{2^(prev-5)}): the polynomial sum_i c^i f_i(y) regrouped as i = j + 32 l. A
// block handles GB = 256/nB consecutive groups g: its 256 lanes first fold the
// ratio-32 groups h = l*2^cur + g (l < nB) of the intermediate layer into LDS,
// then GB lanes fold each g's nB intermediate values. Every read of pol comes
// before the block's writes of out[g], and other blocks read only other g's
// groups, so out may equal pol (in place, as the reference folds).
template <int LOGB>
__global__ __launch_bounds__(256) void k_fri_fold_big(uint64_t *__restrict__ out, const uint64_t *pol,
                                                      uint32_t cur_bits, uint64_t shift_inv1, uint64_t w_inv1,
                                                      uint64_t shift_inv2, uint64_t w_inv2, uint64_t ch0, uint64_t ch1,
                                                      uint64_t ch2, const uint64_t *__restrict__ chp, uint64_t w32_inv,
                                                      uint64_t n32_inv, uint64_t wb_inv, uint64_t nb_inv, RootTab rt1,
                                                      RootTab rt2)
{
    constexpr int NB = 1 << LOGB, GB = 256 / NB;
    __shared__ uint64_t mid[3][256];  // [k][l*GB + gl]
    uint64_t n2 = 1ull << cur_bits, n1 = n2 << LOGB;  // intermediate layer: 2^(cur+LOGB) values
    gl3 c = load_challenge(ch0, ch1, ch2, chp);
    uint32_t t = threadIdx.x, l = t / GB, gl = t % GB;
    uint64_t g = (uint64_t)blockIdx.x * GB + gl;
    if (g < n2) {
        uint64_t h = (uint64_t)l * n2 + g;
        uint64_t a[3][32];
        load_group<5>(a, pol, n1, h);
        gl3 r = fold_group<5>(a, gl_mul(shift_inv1, root_pow(rt1, w_inv1, h)), c, w32_inv, n32_inv);
#pragma unroll
        for (int k = 0; k < 3; k++) mid[k][t] = r.v[k];
    }
    __syncthreads();
    if (t < (uint32_t)GB && g < n2) {
        gl3 c32 = c;
#pragma unroll
        for (int q = 0; q < 5; q++) c32 = gl3_mul(c32, c32);
        uint64_t a[3][NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            int p = (int)bitrev32((uint32_t)i, LOGB);
#pragma unroll
            for (int k = 0; k < 3; k++) a[k][p] = mid[k][i * GB + t];
        }
        gl3 r = fold_group<LOGB>(a, gl_mul(shift_inv2, root_pow(rt2, w_inv2, g)), c32, wb_inv, nb_inv);
#pragma unroll
        for (int k = 0; k < 3; k++) out[g * 3 + k] = r.v[k];
    }
}

__global__ __launch_bounds__(256) void k_fri_transpose(uint64_t *__restrict__ aux, const uint64_t *__restrict__ pol,
                                                       uint64_t degree, uint32_t t_bits)
{
    uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // destination element index i*h + j
    if (d >= degree) return;
    uint64_t w = 1ull << t_bits, h = degree >> t_bits;
    uint64_t i = d / h, j = d - i * h;
    uint64_t s = j * w + i;
#pragma unroll
    for (int k = 0; k < 3; k++) aux[d * 3 + k] = pol[s * 3 + k];
}

extern "C" int fri_launch_fold(uint64_t *out, const uint64_t *pol, const uint64_t *challenge, uint64_t n_bits_ext,
                               uint64_t prev_bits, uint64_t cur_bits, const NttTablesHost *tw, hipStream_t st)
{
    return fri_launch_fold2(out, pol, challenge, nullptr, n_bits_ext, prev_bits, cur_bits, tw, st);
}

// The inverse split tables of the smallest NTT domain (2^24 or 2^28) holding
// 2^log_w points, scaled so that e0 indexes w_{2^log_w}^-e0; none above 2^28.
static RootTab root_tab(const NttTablesHost *tw, uint64_t log_w)
{
    if (!tw || log_w > 28) return {nullptr, nullptr, 0, 0};
    const NttSplitTables &S = tw->split[log_w > 24 ? 1 : 0];
    return {S.twl_i, S.twh_i, S.bits, (uint32_t)(2 * S.bits - log_w)};
}

extern "C" int fri_launch_fold2(uint64_t *out, const uint64_t *pol, const uint64_t *challenge,
                                const uint64_t *d_challenge, uint64_t n_bits_ext, uint64_t prev_bits,
                                uint64_t cur_bits, const NttTablesHost *tw, hipStream_t st)
{
    if (cur_bits > prev_bits || prev_bits > n_bits_ext || prev_bits > 32) return -1;
    uint32_t logx = (uint32_t)(prev_bits - cur_bits);
    uint64_t shift = 7;
    for (uint64_t j = 0; j < n_bits_ext - prev_bits; j++) shift = gl_mul(shift, shift);
    uint64_t shift_inv = gl_inv(shift);
    static const uint64_t W[33] = {
        0x1ULL, 18446744069414584320ULL, 281474976710656ULL, 16777216ULL, 4096ULL, 64ULL, 8ULL,
        2198989700608ULL, 4404853092538523347ULL, 6434636298004421797ULL, 4255134452441852017ULL,
        9113133275150391358ULL, 4355325209153869931ULL, 4308460244895131701ULL,
        7126024226993609386ULL, 1873558160482552414ULL, 8167150655112846419ULL,
        5718075921287398682ULL, 3411401055030829696ULL, 8982441859486529725ULL,
        1971462654193939361ULL, 6553637399136210105ULL, 8124823329697072476ULL,
        5936499541590631774ULL, 2709866199236980323ULL, 8877499657461974390ULL,
        3757607247483852735ULL, 4969973714567017225ULL, 2147253751702802259ULL,
        2530564950562219707ULL, 1905180297017055339ULL, 3524815499551269279ULL,
        7277203076849721926ULL};
    uint64_t w_inv = gl_inv(W[prev_bits]);
    uint64_t c0 = 0, c1 = 0, c2 = 0;
    if (!d_challenge) c0 = gl_canon(challenge[0]), c1 = gl_canon(challenge[1]), c2 = gl_canon(challenge[2]);
    uint64_t n2 = 1ull << cur_bits;
    if (logx > 5) {
        // two-stage fold: ratio 32 (prev -> prev-5) at c, then ratio 2^(logx-5) at c^32
        uint32_t logb = logx - 5;
        uint64_t shift2 = shift;
        for (int q = 0; q < 5; q++) shift2 = gl_mul(shift2, shift2);
        uint64_t groups = 256u >> logb, blocks = (n2 + groups - 1) / groups;
        dim3 gb((unsigned)blocks);
        uint64_t w32_inv = gl_inv(W[5]), n32_inv = gl_inv(32), wb_inv = gl_inv(W[logb]);
        uint64_t nb_inv = gl_inv(1ull << logb), w_inv2 = gl_inv(W[prev_bits - 5]), shift_inv2 = gl_inv(shift2);
#define FOLD_BIG(L)                                                                                              \
    case L:                                                                                                      \
        hipLaunchKernelGGL(k_fri_fold_big<L>, gb, dim3(256), 0, st, out, pol, (uint32_t)cur_bits, shift_inv,     \
                           w_inv, shift_inv2, w_inv2, c0, c1, c2, d_challenge, w32_inv, n32_inv, wb_inv, nb_inv,  \
                           root_tab(tw, prev_bits), root_tab(tw, prev_bits - 5));                                \
        break;
        switch (logb) {
            FOLD_BIG(1)
            FOLD_BIG(2)
            FOLD_BIG(3)
            FOLD_BIG(4)
            FOLD_BIG(5)
        default: return -2;
        }
#undef FOLD_BIG
        return (int)hipGetLastError();
    }
    uint64_t wx_inv = gl_inv(W[logx]);
    uint64_t nx_inv = gl_inv((1ull << logx) % GL_P);
    dim3 g((unsigned)((n2 + 255) / 256));
#define FOLD_CASE(L)                                                                                        \
    case L:                                                                                                 \
        hipLaunchKernelGGL(k_fri_fold<L>, g, dim3(256), 0, st, out, pol, (uint32_t)cur_bits, shift_inv,     \
                           w_inv, c0, c1, c2, d_challenge, wx_inv, nx_inv, root_tab(tw, prev_bits));        \
        break;
    switch (logx) {
        FOLD_CASE(1)
        FOLD_CASE(2)
        FOLD_CASE(3)
        FOLD_CASE(4)
        FOLD_CASE(5)
    default: return -2;
    }
#undef FOLD_CASE
    return (int)hipGetLastError();
}

extern "C" int fri_launch_transpose(uint64_t *aux, const uint64_t *pol, uint64_t degree, uint32_t t_bits,
                                    hipStream_t st)
{
    if (degree == 0) return 0;
    dim3 g((unsigned)((degree + 255) / 256));
    hipLaunchKernelGGL(k_fri_transpose, g, dim3(256), 0, st, aux, pol, degree, t_bits);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- field self-check ----
// Element-wise field operations on device (op 0 add, 1 sub, 2 mul, 3 canonical
// reduce of a 128-bit value (a = low word, b = high word)). Exposed so the GPU
// tests can pin the device arithmetic on edge values against exact integers.
__global__ __launch_bounds__(256) void k_gl_binop(uint64_t *__restrict__ out, const uint64_t *__restrict__ a,
                                                  const uint64_t *__restrict__ b, uint64_t n, int op)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t x = a[i], y = b[i];
    uint64_t r;
    if (op == 0) r = gl_add(gl_canon(x), gl_canon(y));
    else if (op == 1) r = gl_sub(gl_canon(x), gl_canon(y));
    else if (op == 2) r = gl_mul(x, y);
    else r = gl_reduce128_canon(x, y);
    out[i] = r;
}

extern "C" int fri_launch_gl_binop(uint64_t *out, const uint64_t *a, const uint64_t *b, uint64_t n, int op,
                                   hipStream_t st)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gl_binop, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, a, b, n, op);
    return (int)hipGetLastError();
}
