// Goldilocks LDE kernels for gfx950 (MI355X).
//
// Implements the semantics of NTT_Goldilocks::LDE
// (pil2-stark/src/goldilocks/src/ntt_goldilocks.cpp:369-404): for every column,
// out[j] = p(7 * w_NExt^j) where p interpolates the column over <w_N>.
//
// Device layout: a trace arrives row-major (N x C, the CPU path's layout,
// starks.hpp:151-152) and is transposed once into column-major scratch; the LDE
// is produced column-major (column c at lde + c*NExt), which keeps every NTT pass
// and the row hashing coalesced.
//
// Transform plan (four-step, no explicit bit-reversal pass), N = n1*n2, NExt = N*B,
// with n2*B <= 2^14 (one LDS tile) and n1 <= 2^10:
//   pass 1  strided iNTT: n1-point DIF over m (stride n2), then x w_N^-(i*bitrev(m))
//   pass 2  one block per (column, m): n2-point DIF -> coefficients in bit-reversed
//           order, x 7^e/N, spread (in bit-reversed order the zero-padded array
//           holds each coefficient followed by B-1 zeros, and the first log2(B) DIT
//           stages make B copies of it), n2*B-point DIT from stage log2(B) on,
//           then x w_NExt^(bitrev(m)*k)
//   pass 3  strided NTT: n1-point DIT over m (stride n2*B) -> natural order
//
// Inside a tile every sub-transform of 2^L points runs as rounds of at most 4
// index bits (four-step again): a 2^r-point DFT in registers whose internal
// twiddles are powers of w_16, and one twiddle per element between rounds,
// w_{2^(q+r)}^(o*bitrev(j)) (o = the element's index bits below the round).
// Each thread keeps 16 values per round; rounds exchange through LDS. The first
// round of a pass reads global memory and the last one writes it, both in runs
// of >= 64 bytes (whole 128-byte lines per XCD; see strided_tile).
#include <algorithm>

#include "gl_field.h"
#include "ntt_kernels.h"

// --------------------------------------------------------------- tables ----
// All roots derive from the reference's chain W[k] (goldilocks_base_field.cpp:11-44).
//   tkf[2^k + y] = w_{2^k}^y, tki[2^k + y] = w_{2^k}^-y  (k <= 14, y < 2^k)  inside tiles:
//        one compact table per round size, so a round's twiddles of 2^k <= 2^12
//        entries fit the 32 KiB L1 instead of being spread over a 2^14-entry table
//   twl/twh: w_D^x = lo[x & (2^bits - 1)] * hi[x >> bits]   (x < D)   between passes
//   shl/shh: 7^e = lo[e & (2^bits - 1)] * hi[e >> bits]              coset shift
// over the domain D = 2^(2*bits): bits = 12 (D = 2^24, lo tables of 32 KiB) for
// transforms up to 2^24 points, bits = 14 (D = 2^28) above that.
struct NttTables {
    const uint64_t *tkf, *tki;
    const uint64_t *twl_f, *twh_f, *twl_i, *twh_i;
    const uint64_t *shl, *shh;
    uint32_t bits;
};

// the tables for transforms of at most 2^log_max points
static NttTables tables_for(const NttTablesHost *th, uint32_t log_max)
{
    const NttSplitTables &S = th->split[log_max > 24 ? 1 : 0];
    return {th->tkf, th->tki, S.twl_f, S.twh_f, S.twl_i, S.twh_i, S.shl, S.shh, S.bits};
}

__device__ __forceinline__ uint32_t tw_lo(uint32_t x, const NttTables &tb) { return x & ((1u << tb.bits) - 1u); }
__device__ __forceinline__ uint32_t tw_hi(uint32_t x, const NttTables &tb) { return x >> tb.bits; }

__device__ __forceinline__ uint32_t lds_addr(uint32_t x) { return x + (x >> 4); }
// For tile indices with disjoint bits, lds_addr(a | b) = lds_addr(a) + lds_addr(b): a
// round's values sit at lds_addr(base) + lds_addr(j << shift), the second term the
// same for every lane (a scalar offset).

// ------------------------------------------------------------ transpose ----
// in: rows x cols row-major (row stride in_rs), out: column c at out + c*out_cs.
#define TR_T 32
__global__ __launch_bounds__(256) void k_transpose_rm2cm(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                         uint64_t rows, uint64_t cols, uint64_t in_rs, uint64_t out_cs)
{
    __shared__ uint64_t tile[TR_T][TR_T + 1];
    uint64_t r0 = (uint64_t)blockIdx.x * TR_T, c0 = (uint64_t)blockIdx.y * TR_T;
    uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = __builtin_nontemporal_load(&in[r * in_rs + c]);
    }
    __syncthreads();
    for (uint32_t k = ty; k < TR_T; k += 8) {  // loads and stores are non-temporal: each word moves once
        uint64_t c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) __builtin_nontemporal_store(tile[tx][k], &out[c * out_cs + r]);
    }
}

// column-major (col c at in + c*in_cs) -> row-major rows x cols (row stride out_rs)
__global__ __launch_bounds__(256) void k_transpose_cm2rm(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                         uint64_t rows, uint64_t cols, uint64_t in_cs, uint64_t out_rs)
{
    __shared__ uint64_t tile[TR_T][TR_T + 1];
    uint64_t r0 = (uint64_t)blockIdx.x * TR_T, c0 = (uint64_t)blockIdx.y * TR_T;
    uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[c * in_cs + r];
    }
    __syncthreads();
    for (uint32_t k = ty; k < TR_T; k += 8) {
        uint64_t r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) out[r * out_rs + c] = tile[tx][k];
    }
}

// ------------------------------------------------------ small DFTs -------
// R-point DIF (natural in -> bit-reversed out) on v[0..R), roots w_R^k = w_16^(k*16/R).
template <int R> struct Log2 { static constexpr int v = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : 1; };

// Butterfly k of a stage with half-size H (R/2 butterflies) pairs values
// b + jj and b + jj + H, b = (k / H) * 2H, jj = k mod H; its twiddle is
// w_{2H}^jj = w_16^(jj * 8 / H). The m-th butterfly with jj != 0:
template <int H>
__device__ constexpr int tw_k(int m)
{
    return (m / (H - 1)) * H + m % (H - 1) + 1;
}

// The internal twiddles are powers of w_16 = 2^12 (W[4] of the reference's root
// chain): w_16^k = 2^(12k) and w_16^-k = -2^(96 - 12k), since 2^96 = -1. They are
// applied as multiplications by a power of two (gl_mul_pow2), and the sign of
// the inverse roots is folded into the stage's additions and subtractions.
template <bool INV>
__device__ constexpr int w16_shift(int k)
{
    return INV ? 96 - 12 * k : 12 * k;
}

// v[b] *= 2^(shift of butterfly b's twiddle) for the butterflies b = tw_k<H>(m),
// m in [M, NM), three lanes at a time
template <int H, bool INV, int M, int NM>
__device__ __forceinline__ void stage_twiddles(uint64_t *v)
{
    if constexpr (M + 3 <= NM) {
        constexpr int b0 = tw_k<H>(M), b1 = tw_k<H>(M + 1), b2 = tw_k<H>(M + 2);
        uint64_t t[3] = {v[b0], v[b1], v[b2]};
        gl_mul_pow2_3<w16_shift<INV>(b0 % H * 8 / H), w16_shift<INV>(b1 % H * 8 / H),
                      w16_shift<INV>(b2 % H * 8 / H)>(t, t);
        v[b0] = t[0];
        v[b1] = t[1];
        v[b2] = t[2];
        stage_twiddles<H, INV, M + 3, NM>(v);
    } else if constexpr (M < NM) {
        constexpr int b0 = tw_k<H>(M);
        v[b0] = gl_mul_pow2<w16_shift<INV>(b0 % H * 8 / H)>(v[b0]);
        stage_twiddles<H, INV, M + 1, NM>(v);
    }
}

// One DIF stage on R values in registers: (u, x) -> (u + x, (u - x) * w); with an
// inverse root w = -2^s the difference is taken as x - u. The additions,
// subtractions and twiddle products each run lane-interleaved.
template <int R, bool INV, int ST>
__device__ __forceinline__ void dif_stage(uint64_t *v)
{
    constexpr int H = 1 << ST, NB = R / 2;
    uint64_t u[NB], x[NB], sm[NB], df[NB], da[NB], db[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        u[k] = v[(k / H) * 2 * H + k % H];
        x[k] = v[(k / H) * 2 * H + k % H + H];
        bool neg = INV && (k % H) != 0;
        da[k] = neg ? x[k] : u[k];
        db[k] = neg ? u[k] : x[k];
    }
    gl_add_v<NB>(sm, u, x);
    gl_sub_v<NB>(df, da, db);
    if constexpr (H > 1) stage_twiddles<H, INV, 0, NB / H * (H - 1)>(df);
#pragma unroll
    for (int k = 0; k < NB; k++) {
        v[(k / H) * 2 * H + k % H] = sm[k];
        v[(k / H) * 2 * H + k % H + H] = df[k];
    }
}

// One DIT stage: (u, x) -> (u + x*w, u - x*w); with an inverse root w = -2^s,
// t = x*2^s gives (u - t, u + t).
template <int R, bool INV, int ST>
__device__ __forceinline__ void dit_stage(uint64_t *v)
{
    constexpr int H = 1 << ST, NB = R / 2;
    uint64_t u[NB], x[NB], sa[NB], sd[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        u[k] = v[(k / H) * 2 * H + k % H];
        x[k] = v[(k / H) * 2 * H + k % H + H];
    }
    if constexpr (H > 1) stage_twiddles<H, INV, 0, NB / H * (H - 1)>(x);
    gl_add_v<NB>(sa, u, x);
    gl_sub_v<NB>(sd, u, x);
#pragma unroll
    for (int k = 0; k < NB; k++) {
        bool neg = INV && (k % H) != 0;
        v[(k / H) * 2 * H + k % H] = neg ? sd[k] : sa[k];
        v[(k / H) * 2 * H + k % H + H] = neg ? sa[k] : sd[k];
    }
}

template <int R, bool INV, int ST>
__device__ __forceinline__ void dif_stages(uint64_t *v)
{
    dif_stage<R, INV, ST>(v);
    if constexpr (ST > 0) dif_stages<R, INV, ST - 1>(v);
}

template <int R, bool INV, int ST>
__device__ __forceinline__ void dit_stages(uint64_t *v)
{
    dit_stage<R, INV, ST>(v);
    if constexpr (ST + 1 < Log2<R>::v) dit_stages<R, INV, ST + 1>(v);
}

// R-point DIF (natural in -> bit-reversed out), R <= 16
template <int R, bool INV>
__device__ __forceinline__ void dft_dif(uint64_t *v)
{
    dif_stages<R, INV, Log2<R>::v - 1>(v);
}

// R-point DIT (bit-reversed in -> natural out)
template <int R, bool INV>
__device__ __forceinline__ void dft_dit(uint64_t *v)
{
    dit_stages<R, INV, 0>(v);
}

// v[j] *= w[j] for j = 1 .. R-1, lane-interleaved
template <int R>
__device__ __forceinline__ void twiddle_rest(uint64_t *v, const uint64_t *w)
{
    if constexpr (R > 1) gl_mul_v<R - 1>(v + 1, v + 1, w + 1);
}

// Element position of value j of group gidx in a round over index bits [q, q+r)
// of a tile with 2^lt lines (line index fastest). Returns the tile-linear index
// (m << lt) | ii and the pieces the kernels need. A thread with G groups per round
// takes gidx = gg * threads + t, so that for each gg consecutive lanes hold
// consecutive positions (contiguous LDS and global accesses).
struct RoundPos {
    uint32_t ii, o_lo, o_hi;
};
__device__ __forceinline__ RoundPos round_pos(uint32_t gidx, uint32_t lt, uint32_t q)
{
    RoundPos p;
    p.ii = gidx & ((1u << lt) - 1);
    uint32_t rest = gidx >> lt;
    p.o_lo = rest & ((1u << q) - 1);
    p.o_hi = rest >> q;
    return p;
}
__device__ __forceinline__ uint32_t round_m(const RoundPos &p, uint32_t j, uint32_t q, uint32_t r)
{
    return (p.o_hi << (q + r)) | (j << q) | p.o_lo;
}

// --------------------------------------------------- strided passes ------
// Tile: n1 = 2^L points (index m, global stride n2) x 2^lt lines (index ii, global
// offset i0 + ii); column c = blockIdx.y at data + c*cs. The tile holds 2^(L+lt)
// values, 16 per thread, in LDS at tile index x = (m << lt) | ii. A load phase
// and a store phase move the tile in runs of 2^lt >= 8 consecutive words; the
// rounds between them work in LDS only.
//
// DIF rounds (iNTT pass 1): index bits from the top, 2^r-point DIF, then value j
// times w_{2^(q+r)}^-(o_lo * bitrev(j)).
// A tile that covers index bits [qlo, qlo + L) of m (SPLIT: a pass over more
// bits than one tile holds, in two launches) takes its twiddles at the global
// position: round bits q + qlo, lower index bits (o_lo << qlo) | m_lo.
template <bool SPLIT>
__device__ __forceinline__ uint32_t global_o(uint32_t o_lo, uint32_t qlo, uint32_t m_lo)
{
    return SPLIT ? (o_lo << qlo) | m_lo : o_lo;
}

template <int R, bool SPLIT = false>
__device__ __forceinline__ void strided_dif_round(uint64_t *__restrict__ s, uint32_t lt, uint32_t q,
                                                  const NttTables &tb, uint32_t qlo = 0, uint32_t m_lo = 0)
{
    constexpr int LR = Log2<R>::v;
    constexpr int G = 16 / R;
    uint32_t t = threadIdx.x;
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        RoundPos p = round_pos(gg * blockDim.x + t, lt, q);
        uint64_t *sb = s + lds_addr((round_m(p, 0, q, LR) << lt) | p.ii);
        uint64_t v[R];
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = sb[lds_addr((uint32_t)j << (q + lt))];
        dft_dif<R, true>(v);
        uint32_t og = global_o<SPLIT>(p.o_lo, qlo, m_lo);
        if (og) {
            uint64_t w[R];
            uint32_t base = 1u << (q + (SPLIT ? qlo : 0) + LR);
#pragma unroll
            for (int j = 1; j < R; j++) w[j] = tb.tki[base + og * bitrev32(j, LR)];
            twiddle_rest<R>(v, w);
        }
#pragma unroll
        for (int j = 0; j < R; j++) sb[lds_addr((uint32_t)j << (q + lt))] = v[j];
    }
}

// DIT rounds (NTT pass 3): index bits from the bottom; value j times
// w_{2^(q+r)}^(bitrev(j) * o_lo), then the 2^r-point DIT (INV: inverse roots).
template <int R, bool INV = false, bool SPLIT = false>
__device__ __forceinline__ void strided_dit_round(uint64_t *__restrict__ s, uint32_t lt, uint32_t q,
                                                  const NttTables &tb, uint32_t qlo = 0, uint32_t m_lo = 0)
{
    constexpr int LR = Log2<R>::v;
    constexpr int G = 16 / R;
    uint32_t t = threadIdx.x;
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        RoundPos p = round_pos(gg * blockDim.x + t, lt, q);
        uint64_t *sb = s + lds_addr((round_m(p, 0, q, LR) << lt) | p.ii);
        uint64_t v[R];
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = sb[lds_addr((uint32_t)j << (q + lt))];
        uint32_t og = global_o<SPLIT>(p.o_lo, qlo, m_lo);
        if (og) {
            uint64_t w[R];
            uint32_t base = 1u << (q + (SPLIT ? qlo : 0) + LR);
#pragma unroll
            for (int j = 1; j < R; j++) w[j] = (INV ? tb.tki : tb.tkf)[base + og * bitrev32(j, LR)];
            twiddle_rest<R>(v, w);
        }
        dft_dit<R, INV>(v);
#pragma unroll
        for (int j = 0; j < R; j++) sb[lds_addr((uint32_t)j << (q + lt))] = v[j];
    }
}

// Strided tiles with runs of 2^lt = 8 words use half of each 128-byte line; the
// tile pair (2u, 2u+1) shares the lines. Workgroups are dispatched round-robin
// over the 8 XCDs (each with its own L2), so the pair goes to blocks b and b+8,
// which share an XCD and run close together.
__device__ __forceinline__ uint32_t strided_tile(uint32_t bx, uint32_t lt)
{
    if (lt != 3 || (gridDim.x & 15u)) return bx;
    return (bx & ~15u) | ((bx & 7u) << 1) | ((bx >> 3) & 1u);
}

// A strided tile holds 16 values per thread (the launch gives it tile / 16
// threads): thread t moves elements x = t + k * blockDim.x, k < 16. All sixteen
// global loads are issued before the first LDS write (and all sixteen LDS reads
// before the first global store), so a wave keeps sixteen HBM requests in flight
// instead of waiting out one round trip per element.
// CANON: the words may be non-canonical (a pass over caller data); the forward
// pass reads the block pass's canonical output and loads it as is. NT: non-temporal
// loads (measured faster for the inverse pass, slower for the forward pass).
template <bool CANON, bool NT>
__device__ __forceinline__ void strided_load(uint64_t *__restrict__ s, const uint64_t *__restrict__ col, uint64_t n2,
                                             uint64_t i0, uint32_t lt)
{
    uint64_t v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        uint32_t x = threadIdx.x + k * blockDim.x;
        const uint64_t *a = &col[(uint64_t)(x >> lt) * n2 + i0 + (x & ((1u << lt) - 1))];
        v[k] = NT ? __builtin_nontemporal_load(a) : *a;
    }
#pragma unroll
    for (int k = 0; k < 16; k++) s[lds_addr(threadIdx.x + k * blockDim.x)] = CANON ? gl_canon(v[k]) : v[k];
    __syncthreads();
}

// The tile back to its global positions (element x at col + (x >> lt) * stride +
// (x & (2^lt - 1))), sixteen LDS reads first. The thread index passes through an
// empty asm so that the compiler recomputes the sixteen addresses here instead of
// keeping the load phase's ones live in registers across the rounds.
__device__ __forceinline__ void strided_store(const uint64_t *__restrict__ s, uint64_t *__restrict__ col,
                                              uint64_t stride, uint32_t lt)
{
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    uint64_t v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = s[lds_addr(t + k * blockDim.x)];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        uint32_t x = t + k * blockDim.x;
        col[(uint64_t)(x >> lt) * stride + (x & ((1u << lt) - 1))] = v[k];
    }
}

// Which part of m (Ltot bits, global stride n2) a strided block covers: bits
// [qlo, qlo + L) vary inside the tile; the lines i0 .. i0 + 2^lt - 1 and the other
// bits of m come from the block index (line tile fastest). m_lo: bits below qlo,
// m_hi: bits above qlo + L. The tile's element x sits at col + base + (x >> lt) *
// stride + (x & (2^lt - 1)).
struct StridedBlock {
    uint64_t base, stride;
    uint32_t m_lo, m_hi;
};
template <bool SPLIT>
__device__ __forceinline__ StridedBlock strided_block(uint32_t lt, uint32_t log_n2, uint32_t qlo, uint32_t L)
{
    uint32_t bx = strided_tile(blockIdx.x, lt);
    if (!SPLIT) return {(uint64_t)bx << lt, 1ull << log_n2, 0u, 0u};
    uint32_t tiles_log = log_n2 - lt;
    uint64_t i0 = (uint64_t)(bx & ((1u << tiles_log) - 1)) << lt;
    uint32_t other = bx >> tiles_log;
    uint32_t m_lo = other & ((1u << qlo) - 1), m_hi = other >> qlo;
    uint64_t m0 = ((uint64_t)m_hi << (qlo + L)) | m_lo;
    return {(m0 << log_n2) + i0, 1ull << (log_n2 + qlo), m_lo, m_hi};
}

// iNTT pass 1: after the DIF the value at m is the transform at bitrev(m); it is
// stored times w_N^-(i * bitrev_Ltot(m)) (i = the line, N = 2^(Ltot + log_n2)).
// SPLIT: this launch covers index bits [qlo, qlo + L) of the Ltot-bit DIF (the
// top bits first, in an earlier launch); only the launch with qlo = 0 applies the
// final twiddle (final_tw).
template <bool SPLIT = false>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_strided_inv(uint64_t *__restrict__ data, uint64_t cs, uint32_t L,
                                                                 uint32_t log_n2, uint32_t lt, NttTables tb,
                                                                 uint32_t qlo, uint32_t Ltot, uint32_t final_tw)
{
    extern __shared__ uint64_t s[];
    StridedBlock B = strided_block<SPLIT>(lt, log_n2, qlo, L);
    uint64_t *col = data + (uint64_t)blockIdx.y * cs + B.base;
    strided_load<true, true>(s, col, B.stride, 0, lt);
    // the fewest rounds of at most 4 bits, as even as possible (10 bits: 4 + 3 + 3); the
    // strided tile keeps 16 values per thread whatever the plan's radix
    int q = (int)L, nr = ((int)L + 3) / 4, kr = 0;
    while (q > 0) {
        int r = (q + (nr - kr) - 1) / (nr - kr);
        kr++;
        q -= r;
        switch (r) {
        case 4: strided_dif_round<16, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        case 3: strided_dif_round<8, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        case 2: strided_dif_round<4, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        default: strided_dif_round<2, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        }
        __syncthreads();
    }
    if (SPLIT && !final_tw) {
        strided_store(s, col, B.stride, lt);
        return;
    }
    // store: the tile holds 16 values per thread; four at a time, twiddles
    // interleaved, in an unrolled loop so that later groups' loads can be issued early
    uint32_t Lt = SPLIT ? Ltot : L;
    uint32_t shift_n = 2 * tb.bits - (Lt + log_n2);
    uint64_t i0 = SPLIT ? B.base & ((1ull << log_n2) - 1) : B.base;
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));  // addresses computed here, not kept live across the rounds
#pragma unroll
    for (int it = 0; it < 4; it++) {
        uint32_t x0 = t + it * 4 * blockDim.x;
        uint64_t v[4], lo[4], hi[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t x = x0 + k * blockDim.x;
            uint32_t m = SPLIT ? (B.m_hi << L) | (x >> lt) : x >> lt;
            uint32_t e = (uint32_t)((i0 + (x & ((1u << lt) - 1))) * bitrev32(m, Lt)) << shift_n;
            v[k] = s[lds_addr(x)];
            lo[k] = tb.twl_i[tw_lo(e, tb)];
            hi[k] = tb.twh_i[tw_hi(e, tb)];
        }
        gl_mul_v<4>(lo, lo, hi);
        gl_mul_v<4>(v, v, lo);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t x = x0 + k * blockDim.x;
            col[(uint64_t)(x >> lt) * B.stride + (x & ((1u << lt) - 1))] = v[k];
        }
    }
}

// NTT pass 3: n1-point DIT over m (bit-reversed in, natural out). SPLIT: this
// launch covers index bits [qlo, qlo + L) (the low bits first, in an earlier launch).
template <bool INV = false, bool SPLIT = false>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_strided_fwd(uint64_t *__restrict__ data, uint64_t cs, uint32_t L,
                                                                 uint32_t log_n2, uint32_t lt, NttTables tb,
                                                                 uint32_t qlo)
{
    extern __shared__ uint64_t s[];
    StridedBlock B = strided_block<SPLIT>(lt, log_n2, qlo, L);
    uint64_t *col = data + (uint64_t)blockIdx.y * cs + B.base;
    strided_load<false, false>(s, col, B.stride, 0, lt);
    uint32_t q = 0, nr = (L + 3) / 4, kr = 0;  // rounds as in k_ntt_strided_inv
    while (q < L) {
        uint32_t r = ((L - q) + (nr - kr) - 1) / (nr - kr);
        kr++;
        switch (r) {
        case 4: strided_dit_round<16, INV, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        case 3: strided_dit_round<8, INV, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        case 2: strided_dit_round<4, INV, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        default: strided_dit_round<2, INV, SPLIT>(s, lt, q, tb, qlo, B.m_lo); break;
        }
        __syncthreads();
        q += r;
    }
    strided_store(s, col, B.stride, lt);
}

// A block's line of n contiguous words into LDS, canonical: thread t takes words
// t + i * blockDim.x, i < V, with all V global loads issued before its first LDS
// write. The block launches give n = V * blockDim.x for V >= 2 (a line of 2^L2
// words over 2^(L2 + logb - 4) threads, V = 16 >> logb); V = 1 covers logb >= 4,
// where a thread has at most one word. The words are read once, so the loads are
// non-temporal, as are the block pass's output stores (measured: block pass on
// 2^22 x 128 -> 2^24 columns 19.7 -> 19.0 ms; the strided passes, whose paired tiles
// share 128-byte lines through L2, were slower with non-temporal stores).
template <int V>
__device__ __forceinline__ void line_load(uint64_t *__restrict__ s, const uint64_t *__restrict__ src, uint32_t n)
{
    if constexpr (V == 1) {
        if (threadIdx.x < n) s[lds_addr(threadIdx.x)] = gl_canon(src[threadIdx.x]);
    } else {
        uint64_t v[V];
#pragma unroll
        for (int i = 0; i < V; i++) v[i] = __builtin_nontemporal_load(&src[threadIdx.x + i * blockDim.x]);
#pragma unroll
        for (int i = 0; i < V; i++) s[lds_addr(threadIdx.x + i * blockDim.x)] = gl_canon(v[i]);
    }
}
__device__ __forceinline__ void line_load_b(uint64_t *__restrict__ s, const uint64_t *__restrict__ src, uint32_t n,
                                            uint32_t logb)
{
    switch (logb) {
    case 0: line_load<16>(s, src, n); break;
    case 1: line_load<8>(s, src, n); break;
    case 2: line_load<4>(s, src, n); break;
    case 3: line_load<2>(s, src, n); break;
    default: line_load<1>(s, src, n); break;
    }
}

// ------------------------------------------------------ fused LDE block ----
// Block (m = blockIdx.x, column c = blockIdx.y); line of n2 = 2^L2 values, block
// of n2*B/16 threads.
// Phase 1: the n2 values are loaded into LDS, then DIF rounds run with V = 16/B
// values per thread (so every thread works); the last round scales the
// coefficient at position k (index e = bitrev_{L1+L2}(m*n2 + k)) by 7^e / N.
// LINES = 2: two lines of n2 values sit one after the other in LDS (line index at
// bit L2 of the position) and the rounds run over both; a round's bits and twiddles
// are those of the position within its line.
template <int R, int V, int LINES = 1>
__device__ __forceinline__ void block_dif_round(uint64_t *__restrict__ s, uint32_t q, bool last, uint32_t m,
                                                uint32_t L2, uint32_t log_n, uint64_t n_inv, const NttTables &tb)
{
    constexpr int LR = Log2<R>::v;
    constexpr int G = V / R;
    uint32_t t = threadIdx.x;
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        RoundPos p = round_pos(gg * (((uint32_t)LINES << L2) >> Log2<V>::v) + t, 0, q);
        uint64_t *sb = s + lds_addr(round_m(p, 0, q, LR));
        uint64_t v[R];
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = sb[lds_addr((uint32_t)j << q)];
        dft_dif<R, true>(v);
        if (p.o_lo) {
            uint64_t w[R];
#pragma unroll
            for (int j = 1; j < R; j++) w[j] = tb.tki[(1u << (q + LR)) + p.o_lo * bitrev32(j, LR)];
            twiddle_rest<R>(v, w);
        }
        if (last) {  // coefficient k of the line times 7^e / N, e = bitrev(m*n2 + k)
            uint64_t lo[R], hi[R], ni[R];
#pragma unroll
            for (int j = 0; j < R; j++) {
                uint32_t e = bitrev32((m << L2) | (round_m(p, j, q, LR) & ((1u << L2) - 1)), log_n);
                lo[j] = tb.shl[tw_lo(e, tb)];
                hi[j] = tb.shh[tw_hi(e, tb)];
                ni[j] = n_inv;
            }
            gl_mul_v<R>(lo, lo, hi);
            gl_mul_v<R>(v, v, ni);
            gl_mul_v<R>(v, v, lo);
        }
#pragma unroll
        for (int j = 0; j < R; j++) sb[lds_addr((uint32_t)j << q)] = v[j];
    }
}

template <int V, int LINES = 1>
__device__ __forceinline__ void block_phase1(uint64_t *__restrict__ s, uint32_t m, uint32_t L2, uint32_t log_n,
                                             uint64_t n_inv, const NttTables &tb, bool scale, uint32_t rmax)
{
    constexpr int LV = Log2<V>::v;
    bool active = threadIdx.x < (((uint32_t)LINES << L2) >> LV);
    int rm = LV < (int)rmax ? LV : (int)rmax;
    int q = (int)L2;
    while (q > 0) {
        int r = q >= rm ? rm : q;
        q -= r;
        bool last = (q == 0) && scale;
        if (active) {
            switch (r) {
            case 4: if constexpr (V >= 16) block_dif_round<16, V, LINES>(s, q, last, m, L2, log_n, n_inv, tb); break;
            case 3: if constexpr (V >= 8) block_dif_round<8, V, LINES>(s, q, last, m, L2, log_n, n_inv, tb); break;
            case 2: if constexpr (V >= 4) block_dif_round<4, V, LINES>(s, q, last, m, L2, log_n, n_inv, tb); break;
            default: block_dif_round<2, V, LINES>(s, q, last, m, L2, log_n, n_inv, tb); break;
            }
        }
        __syncthreads();
    }
}

// Phase 2: DIT rounds over bits [logb, logb + L2) of the spread array of n2*B
// values. The first round reads coefficient (o_hi << r) | j for every copy
// o_lo < B (the spread), the last one applies w_NExt^(bitrev_L1(m) * P) and
// stores position P of block m.
template <int R, bool INV = false, bool WIDE = false>
__device__ __forceinline__ void block_dit_round(uint64_t *__restrict__ s, uint64_t *__restrict__ dst, uint32_t q,
                                                uint32_t logb, bool first, bool last, uint32_t e1, uint32_t shift_e,
                                                uint32_t l1, const NttTables &tb)
{
    constexpr int LR = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : 1;
    constexpr int G = 16 / R;
    uint32_t t = threadIdx.x;
    uint64_t v[G][R];
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        RoundPos p = round_pos(gg * blockDim.x + t, 0, q);
        // the first round (q = logb) reads position P >> logb = (P_0 >> logb) | j
        uint32_t sh = first ? q - logb : q;
        const uint64_t *sb = s + lds_addr(round_m(p, 0, q, LR) >> (first ? logb : 0));
#pragma unroll
        for (int j = 0; j < R; j++) v[gg][j] = sb[lds_addr((uint32_t)j << sh)];
    }
    if (first) __syncthreads();  // the spread overwrites coefficients other threads still read
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        RoundPos p = round_pos(gg * blockDim.x + t, 0, q);
        if (p.o_lo) {
            uint64_t w[R];
#pragma unroll
            for (int j = 1; j < R; j++) w[j] = (INV ? tb.tki : tb.tkf)[(1u << (q + LR)) + p.o_lo * bitrev32(j, LR)];
            twiddle_rest<R>(v[gg], w);
        }
        dft_dit<R, INV>(v[gg]);
        if (last) {
            // position P = (j << q) | o_lo of block m times w_NExt^(e1 * P) = D * C_j, with
            // D = w_NExt^(e1 * o_lo) from the split tables (once per group) and
            // C_j = w_NExt^(e1 * j * 2^q) = w_{2^(L1+LR)}^(e1 * j), one compact-table entry
            // per j that is the same for the whole block (WIDE: from the split tables,
            // since L1 + LR may exceed the compact tables' 2^14 points)
            const uint64_t *tl = INV ? tb.twl_i : tb.twl_f, *th = INV ? tb.twh_i : tb.twh_f;
            uint32_t ed = (e1 * p.o_lo) << shift_e;
            uint64_t d = gl_mul(tl[tw_lo(ed, tb)], th[tw_hi(ed, tb)]);
            uint64_t tw[R], dd[R], c[R];
            if (!WIDE) {
                const uint64_t *ct = (INV ? tb.tki : tb.tkf) + (1u << (l1 + LR));
#pragma unroll
                for (int j = 1; j < R; j++) c[j] = ct[e1 * (uint32_t)j];
            } else {
                uint64_t cl[R], ch[R];
#pragma unroll
                for (int j = 1; j < R; j++) {
                    uint32_t ec = ((e1 * (uint32_t)j) << q) << shift_e;
                    cl[j] = tl[tw_lo(ec, tb)];
                    ch[j] = th[tw_hi(ec, tb)];
                }
                gl_mul_v<R - 1>(c + 1, cl + 1, ch + 1);
            }
#pragma unroll
            for (int j = 1; j < R; j++) dd[j] = d;
            tw[0] = d;
            gl_mul_v<R - 1>(tw + 1, dd + 1, c + 1);
            gl_mul_v<R>(v[gg], v[gg], tw);
#pragma unroll
            for (int j = 0; j < R; j++) __builtin_nontemporal_store(v[gg][j], &dst[round_m(p, j, q, LR)]);
        } else {
            uint64_t *sb = s + lds_addr(round_m(p, 0, q, LR));
#pragma unroll
            for (int j = 0; j < R; j++) sb[lds_addr((uint32_t)j << q)] = v[gg][j];
        }
    }
}

// COEF_IN: the line already holds coefficients in phase 1's output order
// (position k = coefficient bitrev_{L1+L2}(m*n2 + k), scaled); phase 1 is skipped
// and the block only extends (computeQ's second half, and with logb = 0 the
// block pass of a plain transform). INV: inverse roots in the DIT and twiddles.
// WIDE: L1 > 10 (transforms above 2^24 points), where the last round's per-j
// twiddles come from the split tables (a separate instantiation, so that the
// common one keeps its registers).
template <bool COEF_IN, bool INV = false, bool WIDE = false>
__global__ __launch_bounds__(NTT_THREADS) void k_lde_block(const uint64_t *__restrict__ in, uint64_t cs_in,
                                                           uint64_t *__restrict__ out, uint64_t cs_out, uint32_t L1,
                                                           uint32_t L2, uint32_t logb, uint64_t n_inv, uint32_t rmax,
                                                           NttTables tb)
{
    extern __shared__ uint64_t s[];
    uint32_t m = blockIdx.x;
    const uint64_t *src = in + (uint64_t)blockIdx.y * cs_in + ((uint64_t)m << L2);
    uint32_t log_n = L1 + L2;
    // phase 1: load the line, iNTT DIF over L2 bits with 16/B values per thread
    line_load_b(s, src, 1u << L2, logb);
    __syncthreads();
    if constexpr (!COEF_IN) {
        switch (logb) {
        case 0: block_phase1<16>(s, m, L2, log_n, n_inv, tb, true, rmax); break;
        case 1: block_phase1<8>(s, m, L2, log_n, n_inv, tb, true, rmax); break;
        case 2: block_phase1<4>(s, m, L2, log_n, n_inv, tb, true, rmax); break;
        default: block_phase1<2>(s, m, L2, log_n, n_inv, tb, true, rmax); break;
        }
    }
    bool first = true;
    // phase 2: NTT DIT over bits [logb, logb + L2) of the n2*B spread values
    uint64_t *dst = out + (uint64_t)blockIdx.y * cs_out + ((uint64_t)m << (L2 + logb));
    uint32_t e1 = bitrev32(m, L1);
    uint32_t shift_e = 2 * tb.bits - (log_n + logb);
    uint32_t top = L2 + logb;
    uint32_t qq = logb;
    first = true;
    while (qq < top) {
        uint32_t r = (top - qq) >= rmax ? rmax : (top - qq);
        bool last = (qq + r == top);
        switch (r) {
        case 4: block_dit_round<16, INV, WIDE>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
        case 3: block_dit_round<8, INV, WIDE>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
        case 2: block_dit_round<4, INV, WIDE>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
        default: block_dit_round<2, INV, WIDE>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
        }
        if (!last) __syncthreads();
        first = false;
        qq += r;
    }
}

// The fused block pass for blowup 4 on two columns at once: block (m, c) takes line
// m of columns 2c and 2c + 1. Phase 1 runs on both lines together (8 values per
// thread instead of 4: radix-8 rounds, so half the rounds and barriers per line,
// every thread busy). Phase 2 needs the whole tile for one line, so it runs on
// the first line while each thread keeps four of the second line's coefficients
// in registers; they go back to LDS for the second line's phase 2.
__global__ __launch_bounds__(NTT_THREADS) void k_lde_block2(const uint64_t *__restrict__ in, uint64_t cs_in,
                                                            uint64_t *__restrict__ out, uint64_t cs_out, uint32_t L1,
                                                            uint32_t L2, uint64_t n_inv, uint32_t rmax, NttTables tb)
{
    constexpr uint32_t logb = 2;
    extern __shared__ uint64_t s[];
    const uint32_t m = blockIdx.x, n2 = 1u << L2, t = threadIdx.x, nt = blockDim.x;  // n2 = 4 * nt
    const uint64_t c0 = 2ull * blockIdx.y;
    const uint64_t *src = in + c0 * cs_in + ((uint64_t)m << L2);
    {
        uint64_t v[8];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = __builtin_nontemporal_load(&src[t + i * nt]);
            v[4 + i] = __builtin_nontemporal_load(&src[cs_in + t + i * nt]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s[lds_addr(t + i * nt)] = gl_canon(v[i]);
            s[lds_addr(n2 + t + i * nt)] = gl_canon(v[4 + i]);
        }
    }
    __syncthreads();
    uint32_t log_n = L1 + L2;
    block_phase1<8, 2>(s, m, L2, log_n, n_inv, tb, true, rmax);
    uint64_t keep[4];  // the second line's coefficients during the first line's phase 2
#pragma unroll
    for (int i = 0; i < 4; i++) keep[i] = s[lds_addr(n2 + t + i * nt)];
    __syncthreads();
    uint32_t e1 = bitrev32(m, L1);
    uint32_t shift_e = 2 * tb.bits - (log_n + logb);
    uint32_t top = L2 + logb;
    for (int line = 0; line < 2; line++) {
        if (line == 1) {
            __syncthreads();  // every thread has read the first line's last round from LDS
#pragma unroll
            for (int i = 0; i < 4; i++) s[lds_addr(t + i * nt)] = keep[i];
            __syncthreads();
        }
        uint64_t *dst = out + (c0 + line) * cs_out + ((uint64_t)m << top);
        uint32_t qq = logb;
        bool first = true;
        while (qq < top) {
            uint32_t r = (top - qq) >= rmax ? rmax : (top - qq);
            bool last = (qq + r == top);
            switch (r) {
            case 4: block_dit_round<16, false, false>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
            case 3: block_dit_round<8, false, false>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
            case 2: block_dit_round<4, false, false>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
            default: block_dit_round<2, false, false>(s, dst, qq, logb, first, last, e1, shift_e, L1, tb); break;
            }
            if (!last) __syncthreads();
            first = false;
            qq += r;
        }
    }
}

// --------------------------------------------------------- tiny LDE -------
// N < 16 (a column shorter than one thread's group): one lane per output row j
// of one column evaluates sum_k a_k (7 w_NExt^j)^k after a direct O(N^2) iNTT.
__global__ __launch_bounds__(256) void k_lde_tiny(const uint64_t *__restrict__ in, uint64_t cs_in,
                                                  uint64_t *__restrict__ out, uint64_t cs_out, uint32_t log_n,
                                                  uint32_t log_ext, uint64_t n_inv, uint64_t w_n_inv, uint64_t w_ext)
{
    uint32_t n = 1u << log_n, ne = 1u << log_ext;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    const uint64_t *x = in + (uint64_t)blockIdx.y * cs_in;
    uint64_t pt = gl_mul(7, gl_pow(w_ext, j)), acc = 0, xp = 1;
    for (uint32_t k = 0; k < n; k++) {
        uint64_t a = 0, wk = gl_pow(w_n_inv, k), wp = 1;  // a_k = (1/N) sum_i x_i w_N^-ik
        for (uint32_t i = 0; i < n; i++) {
            a = gl_add(a, gl_mul(gl_canon(x[i]), wp));
            wp = gl_mul(wp, wk);
        }
        acc = gl_add(acc, gl_mul(gl_mul(a, n_inv), xp));
        xp = gl_mul(xp, pt);
    }
    out[(uint64_t)blockIdx.y * cs_out + j] = acc;
}

// ------------------------------------------------------------ computeQ ----
// Semantics of Starks::computeQ (starks.hpp:187-239): q (qDim columns over the
// extended coset) is interpolated over NExt points, coefficient p*N + i of column
// d times shiftIn^p = 7^(-p*N) becomes coefficient i of column p*qDim + d
// (p < qDeg, i < N), and every such column is zero-extended and transformed
// over NExt points.
//
// Plan (NExt = n1 * n2*B, the LDE's split): after the strided inverse pass over
// n1, block m of a q column holds n2*B values; their DIF leaves at tile position
// x = (k << logb) | b the coefficient bitrev_{log NExt}(m*n2*B + x) = p*N + i with
// p = bitrev_logb(b) and i = bitrev_{log N}(m*n2 + k). So block m holds, for every
// p, exactly the n2 coefficients that block m of an N -> NExt LDE of column
// p*qDim + d works on, in the order that block expects: k_q_block_inv writes
// them (times shiftIn^p / NExt) to line m of coefficient column p*qDim + d, and
// k_lde_block<true> plus the strided forward pass finish the extension.
__global__ __launch_bounds__(NTT_THREADS) void k_q_block_inv(const uint64_t *__restrict__ in, uint64_t cs_in,
                                                             uint64_t *__restrict__ coef, uint64_t cs_coef, uint32_t L1,
                                                             uint32_t L2, uint32_t logb, uint32_t q_deg, uint32_t q_dim,
                                                             uint64_t n_inv_ext, uint64_t shift_in, NttTables tb)
{
    extern __shared__ uint64_t s[];
    uint32_t m = blockIdx.x, d = blockIdx.y;
    uint32_t lb = L2 + logb;
    const uint64_t *src = in + (uint64_t)d * cs_in + ((uint64_t)m << lb);
    line_load<16>(s, src, 1u << lb);
    __syncthreads();
    block_phase1<16>(s, m, lb, L1 + lb, 0, tb, false, 4);
    uint32_t n2 = 1u << L2;
    uint64_t sc = n_inv_ext;  // shiftIn^p / NExt
    for (uint32_t p = 0; p < q_deg; p++) {
        uint64_t *dst = coef + (uint64_t)(p * q_dim + d) * cs_coef + ((uint64_t)m << L2);
        uint32_t b = bitrev32(p, logb);
        for (uint32_t k = threadIdx.x; k < n2; k += blockDim.x) dst[k] = gl_mul(s[lds_addr((k << logb) | b)], sc);
        sc = gl_mul(sc, shift_in);
    }
}

// NExt < 16 or N == 1: one lane per output row j of column c = p*qDim + d, by the
// definition: a_i = shiftIn^p / NExt * sum_k q_d[k] w^-((p*N + i)*k), then
// out_j = sum_{i<N} a_i w^(i*j), w = w_NExt. q element (k, d) at q[k*rs + d*cs].
__global__ __launch_bounds__(256) void k_q_tiny(const uint64_t *__restrict__ q, uint64_t rs, uint64_t cs,
                                                uint64_t *__restrict__ out, uint32_t log_n, uint32_t log_ext,
                                                uint32_t q_dim, uint64_t n_inv_ext, uint64_t shift_in, uint64_t w_inv,
                                                uint64_t w)
{
    uint32_t n = 1u << log_n, ne = 1u << log_ext;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    uint32_t c = blockIdx.y, p = c / q_dim, d = c - p * q_dim;
    uint64_t sc = n_inv_ext;
    for (uint32_t t = 0; t < p; t++) sc = gl_mul(sc, shift_in);
    uint64_t wj = gl_pow(w, j), acc = 0, xp = 1;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t wk = gl_pow(w_inv, p * n + i), a = 0, wp = 1;
        for (uint32_t k = 0; k < ne; k++) {
            a = gl_add(a, gl_mul(gl_canon(q[(uint64_t)k * rs + (uint64_t)d * cs]), wp));
            wp = gl_mul(wp, wk);
        }
        acc = gl_add(acc, gl_mul(gl_mul(a, sc), xp));
        xp = gl_mul(xp, wj);
    }
    out[(uint64_t)c * ne + j] = acc;
}

// ------------------------------------------------------- plain NTT / INTT ----
// NTT_Goldilocks::NTT / INTT (ntt_goldilocks.cpp:211-260): natural order in and
// out, the inverse scaled by 1/M. Plan: a bit-reversal pass (which also applies
// 1/M for the inverse) into the order k_lde_block<true> expects with logb = 0,
// then that block pass and the strided DIT pass, both with the transform's roots.
//
// Bit reversal, M = 2^lm >= 2^12: position pos = (h, mid, l) with h the top 6 bits
// and l the low 6 bits takes source index bitrev(pos) = (bitrev6(l),
// bitrev(mid), bitrev6(h)). One block per (mid, column) moves a 64 x 64 tile:
// reads 64 runs of 64 words (rows a = bitrev6(l)), writes 64 runs of 64 words.
#define BR_T 64
__global__ __launch_bounds__(256) void k_bitrev_tile(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                                     uint64_t cs, uint32_t lm, uint64_t scale)
{
    __shared__ uint64_t tile[BR_T * (BR_T + 1)];
    uint32_t mid = blockIdx.x;
    uint64_t hi_stride = 1ull << (lm - 6);
    uint64_t src_mid = (uint64_t)bitrev32(mid, lm - 12) << 6, dst_mid = (uint64_t)mid << 6;
    const uint64_t *s = src + (uint64_t)blockIdx.y * cs;
    uint64_t *d = dst + (uint64_t)blockIdx.y * cs;
    for (uint32_t x = threadIdx.x; x < BR_T * BR_T; x += blockDim.x) {
        uint32_t a = x >> 6, b = x & 63;
        tile[a * (BR_T + 1) + b] = gl_canon(s[a * hi_stride + src_mid + b]);
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < BR_T * BR_T; x += blockDim.x) {
        uint32_t h = x >> 6, l = x & 63;
        uint64_t v = tile[bitrev32(l, 6) * (BR_T + 1) + bitrev32(h, 6)];
        d[h * hi_stride + dst_mid + l] = scale == 1 ? v : gl_mul(v, scale);
    }
}

// M < 2^12: one block per column holds the whole column in LDS.
__global__ __launch_bounds__(256) void k_bitrev_small(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                                      uint64_t cs, uint32_t lm, uint64_t scale)
{
    __shared__ uint64_t col[1u << 11];
    const uint64_t *s = src + (uint64_t)blockIdx.x * cs;
    uint64_t *d = dst + (uint64_t)blockIdx.x * cs;
    uint32_t m = 1u << lm;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) col[i] = gl_canon(s[i]);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        uint64_t v = col[bitrev32(i, lm)];
        d[i] = scale == 1 ? v : gl_mul(v, scale);
    }
}

// M < 16: one lane per column, out_k = scale * sum_j x_j w^(j*k), element (r, c)
// at base[r*rs + c*cs]; the column is read completely before it is written, so
// dst may equal src.
__global__ __launch_bounds__(256) void k_ntt_tiny(const uint64_t *src, uint64_t *dst, uint64_t rs, uint64_t cs,
                                                  uint64_t ncols, uint32_t lm, uint64_t w, uint64_t scale)
{
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    uint32_t m = 1u << lm;
    uint64_t x[8], y[8];
    for (uint32_t j = 0; j < m; j++) x[j] = gl_canon(src[j * rs + c * cs]);
    uint64_t wk = 1;
    for (uint32_t k = 0; k < m; k++) {
        uint64_t acc = 0, p = 1;
        for (uint32_t j = 0; j < m; j++) {
            acc = gl_add(acc, gl_mul(x[j], p));
            p = gl_mul(p, wk);
        }
        y[k] = gl_mul(acc, scale);
        wk = gl_mul(wk, w);
    }
    for (uint32_t k = 0; k < m; k++) dst[k * rs + c * cs] = y[k];
}

// ------------------------------------------------------------ launchers ----
static constexpr uint64_t kW[33] = {
    0x1ULL, 18446744069414584320ULL, 281474976710656ULL, 16777216ULL, 4096ULL, 64ULL, 8ULL,
    2198989700608ULL, 4404853092538523347ULL, 6434636298004421797ULL, 4255134452441852017ULL,
    9113133275150391358ULL, 4355325209153869931ULL, 4308460244895131701ULL,
    7126024226993609386ULL, 1873558160482552414ULL, 8167150655112846419ULL,
    5718075921287398682ULL, 3411401055030829696ULL, 8982441859486529725ULL,
    1971462654193939361ULL, 6553637399136210105ULL, 8124823329697072476ULL,
    5936499541590631774ULL, 2709866199236980323ULL, 8877499657461974390ULL,
    3757607247483852735ULL, 4969973714567017225ULL, 2147253751702802259ULL,
    2530564950562219707ULL, 1905180297017055339ULL, 3524815499551269279ULL,
    7277203076849721926ULL};  // goldilocks_base_field.cpp:11-44

static bool g_attr_done = false;

extern "C" int ntt_init_device(void)
{
    // the register DFTs multiply by powers of two: w_16 must be 2^12
    static_assert(kW[4] == 4096ull, "w_16 of the root chain is 2^12");
    hipError_t e = hipSuccess;
    // > 64 KiB of dynamic LDS
    if (!g_attr_done) {
        int bytes = (int)(NTT_LDS_WORDS * 8);
        const void *strided[] = {(const void *)k_ntt_strided_inv<false>, (const void *)k_ntt_strided_inv<true>,
                                 (const void *)k_ntt_strided_fwd<false, false>,
                                 (const void *)k_ntt_strided_fwd<true, false>,
                                 (const void *)k_ntt_strided_fwd<false, true>,
                                 (const void *)k_ntt_strided_fwd<true, true>};
        for (const void *f : strided)
            if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)k_q_block_inv, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        const void *blocks[] = {(const void *)k_lde_block2, (const void *)k_lde_block<false, false, false>,
                                (const void *)k_lde_block<false, false, true>,
                                (const void *)k_lde_block<true, false, false>,
                                (const void *)k_lde_block<true, false, true>,
                                (const void *)k_lde_block<true, true, false>,
                                (const void *)k_lde_block<true, true, true>};
        for (const void *f : blocks)
            if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return (int)e;
        g_attr_done = true;
    }
    return 0;
}

// Strided-pass tile of the computeQ and plain-NTT paths (the LDE takes it from its plan).
#define NTT_STRIDED_TILE 13

static size_t lds_bytes(uint32_t log_tile) { return (size_t)8 * ((1u << log_tile) + (1u << log_tile) / 16); }

// Strided-pass tile: n1 = 2^L1 points x 2^lt lines with 2^(L1 + lt) <= 2^tile values and
// runs of 2^lt >= 8 words (a tile below that runs at the 2^14 maximum instead). Runs of
// 8 words (half lines, tile pairs on one XCD) let a 2^13 tile hold n1 = 2^10 points, so
// two blocks share a CU and one block's loads overlap the other's rounds.
static uint32_t strided_lt(uint32_t L1, uint32_t lines_log, uint32_t tile)
{
    uint32_t lt = tile >= L1 + 3 ? tile - L1 : NTT_LOG_TILE - L1;
    return lt > lines_log ? lines_log : lt;
}

// A strided pass over L1 index bits of m (lines 2^lines_log, stride 2^lines_log
// per m): one launch when L1 <= NTT_MAX_STRIDED, otherwise two, each over about
// half the bits (the DIF of the inverse pass takes the top bits first, the DIT of
// the forward pass the low bits first). Part k covers bits [qlo, qlo + L) with
// runs of 2^lt words.
struct StridedPart {
    uint32_t L, qlo, lt;
};
static int strided_parts(uint32_t L1, uint32_t lines_log, uint32_t tile, bool top_first, StridedPart *part)
{
    if (L1 == 0) return 0;
    if (L1 <= NTT_MAX_STRIDED) {
        part[0] = {L1, 0, strided_lt(L1, lines_log, tile)};
        return 1;
    }
    uint32_t lo = L1 / 2, hi = L1 - lo;  // bits [0, lo) and [lo, L1)
    StridedPart a = {hi, lo, strided_lt(hi, lines_log, tile)}, b = {lo, 0, strided_lt(lo, lines_log, tile)};
    part[0] = top_first ? a : b;
    part[1] = top_first ? b : a;
    return 2;
}

static void launch_strided_inv(uint64_t *data, uint64_t cs, uint32_t L1, uint32_t log_n2, uint32_t tile,
                               uint64_t ncols, const NttTables &tb, hipStream_t st)
{
    StridedPart part[2];
    int np = strided_parts(L1, log_n2, tile, true, part);
    for (int k = 0; k < np; k++) {
        const StridedPart &P = part[k];
        uint32_t log_tile = P.L + P.lt;
        dim3 g((unsigned)((1ull << (log_n2 - P.lt)) << (L1 - P.L)), (unsigned)ncols);
        if (np == 1)
            hipLaunchKernelGGL(k_ntt_strided_inv<false>, g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st, data,
                               cs, P.L, log_n2, P.lt, tb, 0u, L1, 1u);
        else
            hipLaunchKernelGGL(k_ntt_strided_inv<true>, g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st, data,
                               cs, P.L, log_n2, P.lt, tb, P.qlo, L1, (uint32_t)(k == np - 1));
    }
}

template <bool INV>
static void launch_strided_fwd(uint64_t *data, uint64_t cs, uint32_t L1, uint32_t log_n2, uint32_t tile,
                               uint64_t ncols, const NttTables &tb, hipStream_t st)
{
    StridedPart part[2];
    int np = strided_parts(L1, log_n2, tile, false, part);
    for (int k = 0; k < np; k++) {
        const StridedPart &P = part[k];
        uint32_t log_tile = P.L + P.lt;
        dim3 g((unsigned)((1ull << (log_n2 - P.lt)) << (L1 - P.L)), (unsigned)ncols);
        if (np == 1)
            hipLaunchKernelGGL((k_ntt_strided_fwd<INV, false>), g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st,
                               data, cs, P.L, log_n2, P.lt, tb, 0u);
        else
            hipLaunchKernelGGL((k_ntt_strided_fwd<INV, true>), g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st,
                               data, cs, P.L, log_n2, P.lt, tb, P.qlo);
    }
}

// Launch geometry of an LDE of 2^log_n -> 2^log_ext rows under a plan: n2 * B <=
// 2^tile values per block (tile <= 14), n1 = N / n2 <= 2^10 for one strided launch;
// a shape that needs n1 > 2^10 at the requested tile gets the smallest tile that
// fits, and beyond 2^14 values per block the strided passes take two launches
// (n1 <= 2^20). Returns nonzero for a shape no plan covers (log_n >= 4 assumed).
struct LdeGeometry {
    uint32_t tile, rmax, L1, L2, lt_inv, lt_fwd;
};
static int lde_geometry(uint32_t log_n, uint32_t log_ext, const NttPlan *plan, LdeGeometry *g)
{
    uint32_t log_b = log_ext - log_n;
    uint32_t tile = plan ? plan->tile_log : NTT_LOG_TILE, rmax = plan ? plan->radix_log : 4;
    if (tile > NTT_LOG_TILE) tile = NTT_LOG_TILE;
    if (tile < log_b + 4) tile = log_b + 4;
    if (rmax < 1 || rmax > 4) rmax = 4;
    uint32_t L2 = log_n;
    if (L2 + log_b > tile) L2 = tile - log_b;
    if (log_n - L2 > NTT_MAX_STRIDED) L2 = std::min(log_n - NTT_MAX_STRIDED, (uint32_t)NTT_LOG_TILE - log_b);
    uint32_t L1 = log_n - L2;
    if (L2 + log_b > NTT_LOG_TILE || L2 < 4 || L1 > 2 * NTT_MAX_STRIDED) return -2;  // n2*B <= 2^14, >= 16 values
    StridedPart pi[2], pf[2];
    strided_parts(L1, L2, tile, true, pi);
    strided_parts(L1, L2 + log_b, tile, false, pf);
    *g = {tile, rmax, L1, L2, L1 ? pi[0].lt : 0, L1 ? pf[0].lt : 0};
    return 0;
}

extern "C" uint64_t ntt_lde_plan_signature(uint32_t log_n, uint32_t log_ext, const NttPlan *plan)
{
    LdeGeometry g;
    if (log_n < 4 || log_ext < log_n || log_ext > NTT_MAX_LOG || lde_geometry(log_n, log_ext, plan, &g)) return 0;
    // every launch parameter the plan influences (the tile only through L1/L2/lt)
    return 1 + (g.rmax | (uint64_t)g.L2 << 8 | (uint64_t)g.lt_inv << 16 | (uint64_t)g.lt_fwd << 24);
}

extern "C" int ntt_launch_lde(uint64_t *lde, const uint64_t *trace, uint64_t *scratch, uint32_t log_n, uint32_t log_ext,
                              uint64_t ncols, const NttTablesHost *th, const NttPlan *plan, hipStream_t st,
                              hipEvent_t *ev)
{
    // ev (optional): events recorded after passes 0..3 (ev[0..3]) for phase timing
    NttTables tb = tables_for(th, log_ext);
    uint64_t N = 1ull << log_n;
    if (log_ext < log_n || log_ext > NTT_MAX_LOG) return -3;
    uint32_t log_b = log_ext - log_n;
    if (log_b > NTT_LOG_TILE) return -2;
    uint64_t n_inv = gl_inv(N % GL_P);

    // 0. row-major trace -> column-major scratch
    dim3 tg((unsigned)((N + TR_T - 1) / TR_T), (unsigned)((ncols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_rm2cm, tg, dim3(256), 0, st, trace, scratch, N, ncols, ncols, N);
    if (ev) (void)hipEventRecord(ev[0], st);

    if (log_n < 4) {  // columns shorter than 16 rows
        uint64_t ne = 1ull << log_ext;
        hipLaunchKernelGGL(k_lde_tiny, dim3((unsigned)((ne + 255) / 256), (unsigned)ncols), dim3(256), 0, st, scratch,
                           N, lde, ne, log_n, log_ext, n_inv, gl_inv(kW[log_n]), kW[log_ext]);
        for (int i = 1; i < 4; i++)
            if (ev) (void)hipEventRecord(ev[i], st);
        return (int)hipGetLastError();
    }
    LdeGeometry geo;
    if (lde_geometry(log_n, log_ext, plan, &geo)) return -2;
    uint32_t rmax = geo.rmax, L1 = geo.L1, L2 = geo.L2;

    // 1. strided inverse DIF over n1
    launch_strided_inv(scratch, N, L1, L2, geo.tile, ncols, tb, st);
    if (ev) (void)hipEventRecord(ev[1], st);
    // 2. fused block: DIF, coset scale, spread, DIT, twiddle
    {
        uint32_t log_tile = L2 + log_b;
        dim3 g((unsigned)(1u << L1), (unsigned)ncols);
        if (L1 > NTT_MAX_STRIDED) {
            hipLaunchKernelGGL((k_lde_block<false, false, true>), g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile),
                               st, scratch, N, lde, N << log_b, L1, L2, log_b, n_inv, rmax, tb);
        } else if (log_b == 2 && ncols >= 2) {
            // column pairs two lines per block; an odd last column on its own
            dim3 g2((unsigned)(1u << L1), (unsigned)(ncols / 2));
            hipLaunchKernelGGL(k_lde_block2, g2, dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st, scratch, N, lde,
                               N << log_b, L1, L2, n_inv, rmax, tb);
            if (ncols & 1)
                hipLaunchKernelGGL((k_lde_block<false, false, false>), dim3((unsigned)(1u << L1), 1u),
                                   dim3(1u << (log_tile - 4)), lds_bytes(log_tile), st, scratch + (ncols - 1) * N, N,
                                   lde + (ncols - 1) * (N << log_b), N << log_b, L1, L2, log_b, n_inv, rmax, tb);
        } else {
            hipLaunchKernelGGL((k_lde_block<false, false, false>), g, dim3(1u << (log_tile - 4)), lds_bytes(log_tile),
                               st, scratch, N, lde, N << log_b, L1, L2, log_b, n_inv, rmax, tb);
        }
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    // 3. strided forward DIT over n1
    launch_strided_fwd<false>(lde, N << log_b, L1, L2 + log_b, geo.tile, ncols, tb, st);
    if (ev) (void)hipEventRecord(ev[3], st);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_compute_q(uint64_t *cmq, const uint64_t *q, uint64_t q_rs, uint64_t q_cs, uint64_t *coef,
                                    uint32_t log_n, uint32_t log_ext, uint32_t q_deg, uint32_t q_dim,
                                    const NttTablesHost *th, hipStream_t st)
{
    NttTables tb = tables_for(th, log_ext);
    if (log_ext < log_n || log_ext > NTT_MAX_LOG) return -3;
    uint32_t log_b = log_ext - log_n;
    if (q_deg == 0 || q_dim == 0 || q_deg > (1u << log_b)) return -3;
    uint64_t N = 1ull << log_n, NExt = 1ull << log_ext;
    uint32_t ncols = q_deg * q_dim;
    uint64_t n_inv_ext = gl_inv(NExt % GL_P);
    uint64_t shift_in = gl_pow(gl_inv(7), N);  // Goldilocks::shift() = 7
    if (log_n == 0 || log_ext < 4) {
        if (log_ext > 12) return -2;
        hipLaunchKernelGGL(k_q_tiny, dim3((unsigned)((NExt + 255) / 256), ncols), dim3(256), 0, st, q, q_rs, q_cs, cmq,
                           log_n, log_ext, q_dim, n_inv_ext, shift_in, gl_inv(kW[log_ext]), kW[log_ext]);
        return (int)hipGetLastError();
    }
    if (log_b > NTT_LOG_TILE) return -2;
    // the LDE's split: n2*B <= 2^14 values per block, n1 = N / n2 <= 2^10
    uint32_t L2 = log_n;
    if (L2 + log_b > NTT_LOG_TILE) L2 = NTT_LOG_TILE - log_b;
    uint32_t L1 = log_n - L2, lb = L2 + log_b;
    if (L1 > 2 * NTT_MAX_STRIDED || L2 < 1 || lb < 4) return -2;

    // 0. the q columns, column-major, into the last qDim columns of cmQ (consumed
    //    before the extension writes there)
    uint64_t *qc = cmq + (uint64_t)(ncols - q_dim) * NExt;
    if (q_rs == 1) {
        hipError_t e = hipMemcpy2DAsync(qc, NExt * 8, q, q_cs * 8, NExt * 8, q_dim, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return (int)e;
    } else {
        dim3 tg((unsigned)((NExt + TR_T - 1) / TR_T), (unsigned)((q_dim + TR_T - 1) / TR_T));
        hipLaunchKernelGGL(k_transpose_rm2cm, tg, dim3(256), 0, st, q, qc, NExt, (uint64_t)q_dim, q_rs, NExt);
    }
    // 1. strided inverse DIF over n1 (stride n2*B)
    launch_strided_inv(qc, NExt, L1, lb, NTT_STRIDED_TILE, q_dim, tb, st);
    // 2a. inverse blocks -> scaled coefficient lines (coef: column c at coef + c*N)
    hipLaunchKernelGGL(k_q_block_inv, dim3(1u << L1, q_dim), dim3(1u << (lb - 4)), lds_bytes(lb), st, qc, NExt, coef,
                       N, L1, L2, log_b, q_deg, q_dim, n_inv_ext, shift_in, tb);
    // 2b. extension blocks (spread, DIT, twiddle) into cmQ
    if (L1 > NTT_MAX_STRIDED)
        hipLaunchKernelGGL((k_lde_block<true, false, true>), dim3(1u << L1, ncols), dim3(1u << (lb - 4)), lds_bytes(lb),
                           st, coef, N, cmq, NExt, L1, L2, log_b, (uint64_t)0, 4u, tb);
    else
        hipLaunchKernelGGL((k_lde_block<true, false, false>), dim3(1u << L1, ncols), dim3(1u << (lb - 4)),
                           lds_bytes(lb), st, coef, N, cmq, NExt, L1, L2, log_b, (uint64_t)0, 4u, tb);
    // 3. strided forward DIT over n1
    launch_strided_fwd<false>(cmq, NExt, L1, lb, NTT_STRIDED_TILE, ncols, tb, st);
    return (int)hipGetLastError();
}

// M < 16, any layout (element (r, c) at base[r*rs + c*cs]); dst may equal src
extern "C" int ntt_launch_ntt_tiny(uint64_t *dst, const uint64_t *src, uint64_t rs, uint64_t cs, uint32_t lm,
                                   uint64_t ncols, int inverse, hipStream_t st)
{
    if (lm >= 4) return -3;
    uint64_t w = inverse ? gl_inv(kW[lm]) : kW[lm];
    uint64_t scale = inverse ? gl_inv(1ull << lm) : 1;
    hipLaunchKernelGGL(k_ntt_tiny, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, st, src, dst, rs, cs, ncols,
                       lm, w, scale);
    return (int)hipGetLastError();
}

// M >= 16, column-major (column c at base + c*M), dst != src
extern "C" int ntt_launch_ntt(uint64_t *dst, const uint64_t *src, uint32_t lm, uint64_t ncols, int inverse,
                              const NttTablesHost *th, hipStream_t st)
{
    NttTables tb = tables_for(th, lm);
    if (lm > NTT_MAX_LOG) return -3;
    uint64_t M = 1ull << lm;
    uint64_t scale = inverse ? gl_inv(M % GL_P) : 1;
    if (lm < 4 || src == dst) return -3;  // column-major, out of place, M >= 16
    // 1. bit reversal (x 1/M for the inverse)
    if (lm >= 12)
        hipLaunchKernelGGL(k_bitrev_tile, dim3(1u << (lm - 12), (unsigned)ncols), dim3(256), 0, st, src, dst, M, lm,
                           scale);
    else
        hipLaunchKernelGGL(k_bitrev_small, dim3((unsigned)ncols), dim3(256), 0, st, src, dst, M, lm, scale);
    // 2. block DIT over the low L2 bits, then 3. strided DIT over n1
    uint32_t L2 = lm < NTT_LOG_TILE ? lm : NTT_LOG_TILE, L1 = lm - L2;
    if (L1 > 2 * NTT_MAX_STRIDED) return -2;
    dim3 g(1u << L1, (unsigned)ncols);
    if (inverse)
        hipLaunchKernelGGL((L1 > NTT_MAX_STRIDED ? k_lde_block<true, true, true> : k_lde_block<true, true, false>), g,
                           dim3(1u << (L2 - 4)), lds_bytes(L2), st, dst, M, dst, M, L1, L2, 0u, (uint64_t)0, 4u, tb);
    else
        hipLaunchKernelGGL((L1 > NTT_MAX_STRIDED ? k_lde_block<true, false, true> : k_lde_block<true, false, false>), g,
                           dim3(1u << (L2 - 4)), lds_bytes(L2), st, dst, M, dst, M, L1, L2, 0u, (uint64_t)0, 4u, tb);
    if (inverse)
        launch_strided_fwd<true>(dst, M, L1, L2, NTT_STRIDED_TILE, ncols, tb, st);
    else
        launch_strided_fwd<false>(dst, M, L1, L2, NTT_STRIDED_TILE, ncols, tb, st);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_cm2rm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_cs,
                                uint64_t out_rs, hipStream_t st)
{
    dim3 tg((unsigned)((rows + TR_T - 1) / TR_T), (unsigned)((cols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_cm2rm, tg, dim3(256), 0, st, in, out, rows, cols, in_cs, out_rs);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_rm2cm(uint64_t *out, const uint64_t *in, uint64_t rows, uint64_t cols, uint64_t in_rs,
                                uint64_t out_cs, hipStream_t st)
{
    dim3 tg((unsigned)((rows + TR_T - 1) / TR_T), (unsigned)((cols + TR_T - 1) / TR_T));
    hipLaunchKernelGGL(k_transpose_rm2cm, tg, dim3(256), 0, st, in, out, rows, cols, in_rs, out_cs);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- synthetic data ----
// Seeded uniform values in [0, p): element i = mix(seed + (i+1)*0x9E3779B97F4A7C15),
// re-mixed while >= p (counter-based SplitMix64; the numpy generator matches).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_fill_uniform(uint64_t *__restrict__ out, uint64_t n, uint64_t seed,
                                                      uint64_t start)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t z = mix64(seed + (start + i + 1) * 0x9E3779B97F4A7C15ULL);
        while (z >= GL_P) z = mix64(z ^ 0xD1B54A32D192ED03ULL);
        out[i] = z;
    }
}

// Rows of a column group read from page-locked host memory by the kernel itself
// (over PCIe): dst (rows x w, row-major, device) <- columns [0, w) of the host
// rows at src (row stride ld words). 16-byte loads and stores when every piece
// is 16-byte aligned, 8-byte ones otherwise.
__global__ __launch_bounds__(256) void k_gather_rows(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src,
                                                     uint64_t rows, uint32_t w, uint64_t ld, uint32_t wide)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (wide) {
        uint32_t cpr = w / 2;  // 16-byte pieces per row
        uint64_t total = rows * cpr;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
            uint64_t r = i / cpr, c = (i - r * cpr) * 2;
            *(uint4 *)(dst + r * w + c) = *(const uint4 *)(src + r * ld + c);
        }
    } else {
        uint64_t total = rows * w;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
            uint64_t r = i / w, c = i - r * w;
            dst[r * w + c] = src[r * ld + c];
        }
    }
}

extern "C" int ntt_launch_gather_rows(uint64_t *dst, const uint64_t *src, uint64_t rows, uint32_t w, uint64_t ld,
                                      hipStream_t st)
{
    if (rows == 0 || w == 0) return 0;
    uint32_t wide = (w % 2 == 0 && ld % 2 == 0 && ((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 15) == 0) ? 1u : 0u;
    uint64_t items = wide ? rows * (w / 2) : rows * w, blocks = (items + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, rows, w, ld, wide);
    return (int)hipGetLastError();
}

extern "C" int ntt_launch_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed, uint64_t start, hipStream_t st)
{
    if (n == 0) return 0;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)blocks), dim3(256), 0, st, out, n, seed, start);
    return (int)hipGetLastError();
}
