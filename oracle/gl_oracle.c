/*
 * ORACLE — CPU restatement of pil2-stark's polynomial-commitment path.
 *
 * This file is TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker (or as the
 * timed CPU baseline). The product (libpil2_commit_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the pil2-proofman repository root). The restatement is
 * written from the mathematics of each routine, not from the reference's loop
 * structure:
 *   - field:      pil2-stark/src/goldilocks/src/goldilocks_base_field_scalar.hpp:16-103,229-260
 *                 pil2-stark/src/goldilocks/src/goldilocks_base_field.cpp:11-44 (roots W[33])
 *   - cubic ext:  pil2-stark/src/goldilocks/src/goldilocks_cubic_extension.hpp:232-251
 *   - NTT / LDE:  pil2-stark/src/goldilocks/src/ntt_goldilocks.cpp:369-404 (LDE),
 *                 ntt_goldilocks.hpp:44-60 (r_ = 7^i / N)
 *   - Poseidon2:  pil2-stark/src/goldilocks/src/poseidon2_goldilocks.cpp:10-77,130-175;
 *                 poseidon2_goldilocks.hpp:148-240
 *   - tree size:  pil2-stark/src/goldilocks/src/goldilocks_tooling.hpp:9-21
 *   - openings:   pil2-stark/src/starkpil/merkleTree/merkleTreeGL.cpp:84-107,134-175
 *   - FRI:        pil2-stark/src/starkpil/fri/fri.hpp:32-99,178-226
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   gl_mul           : 128-bit '%' reduction on random pairs; W[k]^2 == W[k-1] chain
 *   ntt / intt       : O(n^2) DFT brute force; round trip
 *   lde              : reference golden vectors (test_ntt_cpu.cpp:351-412) and the
 *                      closed form p(7*w^j) (test_ntt_cpu.cpp:415-471); compiled reference
 *   p2_permute       : reference goldens (tests/test_helpers.hpp:22-51), W=8/12/16
 *   p2_linear_hash   : reference goldens (test_helpers.hpp:54-59); size-0 edge
 *   p2_merkletree    : compiled reference (oracle/_ref) roots on random shapes;
 *                      nrows==1 -> linear hash (test_poseidon2_cpu.cpp:418-442)
 *   cubic mul        : associativity / x^3 = x + 1 identity / compiled reference
 *   fri_fold         : per-index brute force (interpolate each coset, evaluate at the
 *                      challenge) and the verify_fold formula (fri.hpp:229-256)
 *   merkle openings  : path recomputation back to the root
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../pil2-proofman_amd/csrc/p2_constants.h"

#define GLP 0xFFFFFFFF00000001ULL

typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- field -- */
/* goldilocks_base_field_tools.hpp:100-105 (toU64 canonicalisation) */
static inline uint64_t o_canon(uint64_t a) { return a >= GLP ? a - GLP : a; }

/* goldilocks_base_field_scalar.hpp:16-42 — all outputs here are canonical. */
static inline uint64_t o_add(uint64_t a, uint64_t b)
{
    a = o_canon(a); b = o_canon(b);
    uint64_t s = a + b;
    if (s < a || s >= GLP) s -= GLP;   /* wrap-around: s - p == s + 2^32 - 1 */
    return s;
}

/* goldilocks_base_field_scalar.hpp:69-103 */
static inline uint64_t o_sub(uint64_t a, uint64_t b)
{
    a = o_canon(a); b = o_canon(b);
    return a >= b ? a - b : a + (GLP - b);
}

/* 2^64 = 2^32 - 1 (mod p) and 2^96 = -1 (mod p): x = lo + 2^64*(h0 + 2^32*h1)
 * = lo + (2^32-1)*h0 - h1.  Published Goldilocks reduction (Plonky2 field). */
static inline uint64_t o_reduce128(u128 x)
{
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t h0 = hi & 0xFFFFFFFFULL, h1 = hi >> 32;
    uint64_t t = o_sub(o_canon(lo), h1);
    uint64_t m = (h0 << 32) - h0;                       /* (2^32-1)*h0 < p */
    return o_add(t, m);
}

/* goldilocks_base_field_scalar.hpp:229-260 */
static inline uint64_t o_mul(uint64_t a, uint64_t b) { return o_reduce128((u128)a * b); }

uint64_t oracle_gl_add(uint64_t a, uint64_t b) { return o_add(a, b); }
uint64_t oracle_gl_sub(uint64_t a, uint64_t b) { return o_sub(a, b); }
uint64_t oracle_gl_mul(uint64_t a, uint64_t b) { return o_mul(a, b); }

static uint64_t o_pow(uint64_t b, uint64_t e)
{
    uint64_t r = 1; b = o_canon(b);
    while (e) { if (e & 1) r = o_mul(r, b); b = o_mul(b, b); e >>= 1; }
    return r;
}
uint64_t oracle_gl_pow(uint64_t b, uint64_t e) { return o_pow(b, e); }

/* goldilocks_base_field.cpp:97-127 (inverse); Fermat: a^(p-2). inv(0) is an error
 * in the reference (throws); here it returns 0 and callers never pass 0. */
static inline uint64_t o_inv(uint64_t a) { return o_pow(a, GLP - 2); }
uint64_t oracle_gl_inv(uint64_t a) { return o_inv(a); }

/* goldilocks_base_field.cpp:11-44: W[k] is a primitive 2^k-th root of unity. */
static const uint64_t O_W[33] = {
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
uint64_t oracle_gl_root(uint32_t k) { return k <= 32 ? O_W[k] : 0; }

/* ------------------------------------------------------- cubic extension -- */
/* F_p[x]/(x^3 - x - 1); goldilocks_cubic_extension.hpp:232-251. Schoolbook
 * product followed by x^3 = x + 1, x^4 = x^2 + x. */
static inline void o_c3mul(uint64_t r[3], const uint64_t a[3], const uint64_t b[3])
{
    uint64_t c0 = o_mul(a[0], b[0]);
    uint64_t c1 = o_add(o_mul(a[0], b[1]), o_mul(a[1], b[0]));
    uint64_t c2 = o_add(o_add(o_mul(a[0], b[2]), o_mul(a[1], b[1])), o_mul(a[2], b[0]));
    uint64_t c3 = o_add(o_mul(a[1], b[2]), o_mul(a[2], b[1]));
    uint64_t c4 = o_mul(a[2], b[2]);
    r[0] = o_add(c0, c3);
    r[1] = o_add(o_add(c1, c3), c4);
    r[2] = o_add(c2, c4);
}
void oracle_c3_mul(uint64_t *r, const uint64_t *a, const uint64_t *b) { uint64_t t[3]; o_c3mul(t, a, b); memcpy(r, t, 24); }

/* ------------------------------------------------------------------ NTT -- */
static uint32_t o_log2(uint64_t n) { uint32_t l = 0; while ((1ULL << l) < n) l++; return l; }

/* In-place natural-order transform of length n = 2^logn with root w:
 * a[k] <- sum_j a[j] * w^(j*k). Bit reversal followed by radix-2 butterflies. */
static void o_dft_inplace(uint64_t *a, uint32_t logn, uint64_t w)
{
    uint64_t n = 1ULL << logn;
    for (uint64_t i = 0, j = 0; i < n; i++) {
        if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
        uint64_t bit = n >> 1;
        while (bit && (j & bit)) { j ^= bit; bit >>= 1; }
        j |= bit;
    }
    if (n == 1) { a[0] = o_canon(a[0]); return; }
    uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2));
    tw[0] = 1;
    for (uint64_t i = 1; i < n / 2; i++) tw[i] = o_mul(tw[i - 1], w);
    for (uint32_t s = 1; s <= logn; s++) {
        uint64_t m = 1ULL << s, h = m >> 1, step = n / m;
        for (uint64_t b = 0; b < n; b += m)
            for (uint64_t j = 0; j < h; j++) {
                uint64_t u = a[b + j], v = o_mul(a[b + j + h], tw[j * step]);
                a[b + j] = o_add(u, v);
                a[b + j + h] = o_sub(u, v);
            }
    }
    free(tw);
}

/* Forward NTT (natural order), ntt_goldilocks.cpp:211-260 semantics. */
void oracle_ntt(uint64_t *a, uint64_t n) { if (n) o_dft_inplace(a, o_log2(n), O_W[o_log2(n)]); }

/* Inverse NTT (natural order, includes 1/n). */
void oracle_intt(uint64_t *a, uint64_t n)
{
    if (!n) return;
    uint32_t l = o_log2(n);
    o_dft_inplace(a, l, o_inv(O_W[l]));
    uint64_t ninv = o_inv(n % GLP);
    for (uint64_t i = 0; i < n; i++) a[i] = o_mul(a[i], ninv);
}

/* Brute-force O(n^2) DFT used only to pin oracle_ntt. */
void oracle_dft_naive(uint64_t *out, const uint64_t *in, uint64_t n, int inverse)
{
    uint32_t l = o_log2(n);
    uint64_t w = inverse ? o_inv(O_W[l]) : O_W[l];
    uint64_t ninv = inverse ? o_inv(n) : 1;
    for (uint64_t k = 0; k < n; k++) {
        uint64_t acc = 0, wk = o_pow(w, k), x = 1;
        for (uint64_t j = 0; j < n; j++) { acc = o_add(acc, o_mul(in[j], x)); x = o_mul(x, wk); }
        out[k] = o_mul(acc, ninv);
    }
}

/* Evaluate polynomial (coeffs c[0..n)) at x (Horner). */
uint64_t oracle_poly_eval(const uint64_t *c, uint64_t n, uint64_t x)
{
    uint64_t r = 0;
    for (uint64_t k = n; k-- > 0;) r = o_add(o_mul(r, x), c[k]);
    return r;
}

/* LDE: ntt_goldilocks.cpp:369-404. in is row-major N x ncols, out row-major
 * NExt x ncols with out[j][c] = p_c(7 * w_NExt^j), p_c the interpolant of column c
 * over <w_N>. The reference computes INTT(extend) -> x r_[i] = 7^i/N
 * (ntt_goldilocks.hpp:48-60) -> zero-extended NTT; restated here per column. */
void oracle_lde(uint64_t *out, const uint64_t *in, uint64_t NExt, uint64_t N, uint64_t ncols)
{
    if (N == 0 || ncols == 0) return;
    uint32_t lN = o_log2(N), lE = o_log2(NExt);
    uint64_t *shiftpow = (uint64_t *)malloc(sizeof(uint64_t) * N);
    shiftpow[0] = 1;
    for (uint64_t i = 1; i < N; i++) shiftpow[i] = o_mul(shiftpow[i - 1], 7);
    uint64_t winvN = o_inv(O_W[lN]), Ninv = o_inv(N);
    #pragma omp parallel
    {
        uint64_t *col = (uint64_t *)malloc(sizeof(uint64_t) * NExt);
        #pragma omp for schedule(dynamic, 1)
        for (uint64_t c = 0; c < ncols; c++) {
            for (uint64_t i = 0; i < N; i++) col[i] = in[i * ncols + c];
            o_dft_inplace(col, lN, winvN);
            for (uint64_t i = 0; i < N; i++) col[i] = o_mul(o_mul(col[i], Ninv), shiftpow[i]);
            for (uint64_t i = N; i < NExt; i++) col[i] = 0;
            o_dft_inplace(col, lE, O_W[lE]);
            for (uint64_t j = 0; j < NExt; j++) out[j * ncols + c] = col[j];
        }
        free(col);
    }
    free(shiftpow);
}

/* Starks::computeQ, pil2-stark/src/starkpil/starks.hpp:187-239. q: row-major
 * NExt x qDim (the extended-coset evaluations); each column is interpolated over
 * <w_NExt> (the reference's in-place INTT, :196-200), coefficient p*N + i of
 * column d times S[p] = shiftIn^p, shiftIn = (1/7)^N (:202-207), lands in row i,
 * column p*qDim + d of cmQ for i < N (:209-216); rows N..NExt are zero
 * (:218-221) and every cmQ column is transformed over NExt points (:222-226).
 * cmq: row-major NExt x (qDeg*qDim). q is not modified. */
void oracle_compute_q(uint64_t *cmq, const uint64_t *q, uint64_t N, uint64_t NExt, uint64_t qDeg, uint64_t qDim)
{
    if (!N || !NExt || !qDeg || !qDim) return;
    uint32_t lE = o_log2(NExt);
    uint64_t nc = qDeg * qDim;
    uint64_t shift_in = o_pow(o_inv(7), N);
    uint64_t *S = (uint64_t *)malloc(sizeof(uint64_t) * qDeg);
    S[0] = 1;
    for (uint64_t p = 1; p < qDeg; p++) S[p] = o_mul(S[p - 1], shift_in);
    uint64_t winv = o_inv(O_W[lE]), ninv = o_inv(NExt % GLP);
    #pragma omp parallel
    {
        uint64_t *col = (uint64_t *)malloc(sizeof(uint64_t) * NExt);
        uint64_t *out = (uint64_t *)malloc(sizeof(uint64_t) * NExt);
        #pragma omp for schedule(dynamic, 1)
        for (uint64_t d = 0; d < qDim; d++) {
            for (uint64_t k = 0; k < NExt; k++) col[k] = q[k * qDim + d];
            o_dft_inplace(col, lE, winv);
            for (uint64_t k = 0; k < NExt; k++) col[k] = o_mul(col[k], ninv);
            for (uint64_t p = 0; p < qDeg; p++) {
                for (uint64_t i = 0; i < N; i++) out[i] = o_mul(col[p * N + i], S[p]);
                for (uint64_t i = N; i < NExt; i++) out[i] = 0;
                o_dft_inplace(out, lE, O_W[lE]);
                for (uint64_t j = 0; j < NExt; j++) cmq[j * nc + p * qDim + d] = out[j];
            }
        }
        free(col);
        free(out);
    }
    free(S);
}

/* ------------------------------------------------------------- Poseidon2 -- */
static const uint64_t *o_rc(uint32_t W) { return W == 8 ? P2_RC8 : W == 12 ? P2_RC12 : P2_RC16; }
static const uint64_t *o_diag(uint32_t W) { return W == 8 ? P2_DIAG8 : W == 12 ? P2_DIAG12 : P2_DIAG16; }

/* 4x4 block of the external matrix (poseidon2_goldilocks.hpp:186-205):
 * the circulant-like M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]]. */
static void o_m4(uint64_t *x)
{
    static const uint64_t M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
    uint64_t y[4];
    for (int i = 0; i < 4; i++) {
        uint64_t acc = 0;
        for (int j = 0; j < 4; j++) acc = o_add(acc, o_mul(M4[i][j], x[j]));
        y[i] = acc;
    }
    memcpy(x, y, sizeof(y));
}

/* External layer: M4 on each 4-block, then add the sum of same-position entries of
 * all blocks (poseidon2_goldilocks.hpp:208-231). */
static void o_ext(uint64_t *x, uint32_t W)
{
    for (uint32_t i = 0; i < W; i += 4) o_m4(x + i);
    uint64_t s[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < W; i++) s[i & 3] = o_add(s[i & 3], x[i]);
    for (uint32_t i = 0; i < W; i++) x[i] = o_add(x[i], s[i & 3]);
}

static inline uint64_t o_pow7(uint64_t x) { uint64_t x2 = o_mul(x, x); return o_mul(o_mul(x2, x2), o_mul(x2, x)); }

/* poseidon2_goldilocks.cpp:10-41 — ext, 4 full rounds, 22 partial, 4 full. */
void oracle_p2_permute(uint64_t *st, uint32_t W)
{
    const uint64_t *C = o_rc(W), *D = o_diag(W);
    o_ext(st, W);
    for (uint32_t r = 0; r < 4; r++) {
        for (uint32_t i = 0; i < W; i++) st[i] = o_pow7(o_add(st[i], C[r * W + i]));
        o_ext(st, W);
    }
    for (uint32_t r = 0; r < 22; r++) {
        st[0] = o_pow7(o_add(st[0], C[4 * W + r]));
        uint64_t s = 0;
        for (uint32_t i = 0; i < W; i++) s = o_add(s, st[i]);
        for (uint32_t i = 0; i < W; i++) st[i] = o_add(o_mul(st[i], D[i]), s);
    }
    for (uint32_t r = 0; r < 4; r++) {
        for (uint32_t i = 0; i < W; i++) st[i] = o_pow7(o_add(st[i], C[4 * W + 22 + r * W + i]));
        o_ext(st, W);
    }
}

/* Sponge over one row: poseidon2_goldilocks.cpp:43-77. The first block starts
 * with capacity zero; each later block copies state[0..4) into the capacity
 * slots; the rate is zero-padded; size 0 gives a zero digest. */
void oracle_p2_linear_hash(uint64_t *out4, const uint64_t *in, uint64_t size, uint64_t stride, uint32_t W)
{
    uint32_t RATE = W - 4;
    uint64_t st[16];
    if (size == 0) { memset(out4, 0, 32); return; }
    for (uint64_t done = 0; done < size;) {
        if (done == 0) memset(st + RATE, 0, 32);
        else memcpy(st + RATE, st, 32);
        uint64_t n = size - done < RATE ? size - done : RATE;
        for (uint32_t i = 0; i < RATE; i++) st[i] = i < n ? in[(done + i) * stride] : 0;
        oracle_p2_permute(st, W);
        done += n;
    }
    for (int i = 0; i < 4; i++) out4[i] = o_canon(st[i]);
}

/* poseidon2_goldilocks.cpp:177-232 (grinding, width 8): the smallest nonce i with
 * canonical(permute([c0, c1, c2, i, 0, 0, 0, 0])[0]) < 2^(64 - n_bits), tried
 * in increasing order below `limit`; UINT64_MAX if none. The reference's tests
 * (test_poseidon2_cpu.cpp:318-340, 448-465) state the contract this way. */
uint64_t oracle_p2_grinding(const uint64_t *ch, uint32_t n_bits, uint64_t limit)
{
    uint64_t level = (uint64_t)1 << (64 - n_bits);
    for (uint64_t i = 0; i < limit; i++) {
        uint64_t st[8] = {ch[0], ch[1], ch[2], i, 0, 0, 0, 0};
        oracle_p2_permute(st, 8);
        if (o_canon(st[0]) < level) return i;
    }
    return UINT64_MAX;
}

/* TranscriptGL (transcriptGL.hpp:11-58, transcriptGL.cpp:4-119): sponge of width
 * W = 4*arity, rate W-4, capacity 4; outputs canonical. */
typedef struct {
    uint64_t state[16], pending[12], out[16];
    uint32_t pc, oc, W;
} OTranscript;

void oracle_tr_init(OTranscript *t, uint32_t arity)
{
    memset(t, 0, sizeof(*t));
    t->W = 4 * arity;
}

static void o_tr_update(OTranscript *t)   /* _updateState, transcriptGL.cpp:12-45 */
{
    uint32_t rate = t->W - 4;
    uint64_t in[16];
    for (uint32_t i = 0; i < rate; i++) in[i] = i < t->pc ? t->pending[i] : 0;
    for (uint32_t i = 0; i < 4; i++) in[rate + i] = t->state[i];
    oracle_p2_permute(in, t->W);
    for (uint32_t i = 0; i < t->W; i++) t->out[i] = t->state[i] = o_canon(in[i]);
    t->oc = t->W;
    for (uint32_t i = 0; i < rate; i++) t->pending[i] = 0;
    t->pc = 0;
}

void oracle_tr_put(OTranscript *t, const uint64_t *in, uint64_t n)   /* put / _add1, :4-10, :47-56 */
{
    for (uint64_t k = 0; k < n; k++) {
        t->pending[t->pc++] = in[k];
        t->oc = 0;
        if (t->pc == t->W - 4) o_tr_update(t);
    }
}

static uint64_t o_tr_field1(OTranscript *t)   /* getFields1, :81-90 */
{
    if (t->oc == 0) o_tr_update(t);
    uint64_t r = t->out[(t->W - t->oc) % t->W];
    t->oc--;
    return r;
}

void oracle_tr_get_field(OTranscript *t, uint64_t *out3)   /* getField, :58-65 */
{
    for (int i = 0; i < 3; i++) out3[i] = o_tr_field1(t);
}

void oracle_tr_get_state(OTranscript *t, uint64_t *out, uint32_t n)   /* getState, :67-79 */
{
    if (t->pc > 0) o_tr_update(t);
    for (uint32_t i = 0; i < n; i++) out[i] = t->state[i];
}

void oracle_tr_get_permutations(OTranscript *t, uint64_t *res, uint64_t n, uint64_t nbits)   /* :92-118 */
{
    uint64_t nfields = (n * nbits - 1) / 63 + 1, cur = 0, bit = 0;
    uint64_t *f = (uint64_t *)malloc(nfields * 8);
    for (uint64_t i = 0; i < nfields; i++) f[i] = o_tr_field1(t);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t a = 0;
        for (uint64_t j = 0; j < nbits; j++) {
            if ((f[cur] >> bit) & 1) a += (uint64_t)1 << j;
            if (++bit == 63) { bit = 0; cur++; }
        }
        res[i] = a;
    }
    free(f);
}

uint64_t oracle_tr_size(void) { return sizeof(OTranscript); }

/* goldilocks_tooling.hpp:9-21 */
uint64_t oracle_tree_num_elements(uint64_t rows, uint64_t arity)
{
    uint64_t total = rows, level = rows;
    while (level > 1) {
        total += (arity - level % arity) % arity;
        level = (level + arity - 1) / arity;
        total += level;
    }
    return total * 4;
}

/* poseidon2_goldilocks.cpp:130-175 (merkletree_seq): leaf digests of each row,
 * then for each level pad with zero digests to a multiple of arity and compress
 * arity digests into one with the W = 4*arity permutation (truncated to 4).
 * Source element (row r, col c) is src[r*row_stride + c*col_stride]. */
void oracle_p2_merkletree(uint64_t *tree, const uint64_t *src, uint64_t ncols, uint64_t nrows,
                          uint64_t row_stride, uint64_t col_stride, uint64_t arity)
{
    uint32_t W = (uint32_t)(4 * arity);
    if (nrows == 0) return;
    #pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < nrows; r++)
        oracle_p2_linear_hash(&tree[r * 4], &src[r * row_stride], ncols, col_stride, W);
    uint64_t pending = nrows, base = 0;
    while (pending > 1) {
        uint64_t pad = (arity - pending % arity) % arity;
        memset(&tree[(base + pending) * 4], 0, pad * 32);
        uint64_t next = (pending + arity - 1) / arity;
        uint64_t nb = base + pending + pad;
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < next; i++) {
            uint64_t st[16];
            memcpy(st, &tree[(base + i * arity) * 4], W * 8);
            oracle_p2_permute(st, W);
            for (int k = 0; k < 4; k++) tree[(nb + i) * 4 + k] = o_canon(st[k]);
        }
        base = nb;
        pending = next;
    }
}

/* merkleTreeGL.cpp:145-175 (getGroupProof/genMerkleProof): row values followed by,
 * for each level until the level size is <= arity^llv (or 1 when llv == 0), the
 * arity-1 sibling digests of the node on the path. Returns the number of u64 written. */
uint64_t oracle_merkle_open(uint64_t *proof, const uint64_t *tree, const uint64_t *src, uint64_t ncols,
                            uint64_t nrows, uint64_t row_stride, uint64_t col_stride,
                            uint64_t arity, uint64_t llv, uint64_t idx)
{
    uint64_t w = 0;
    for (uint64_t c = 0; c < ncols; c++) proof[w++] = o_canon(src[idx * row_stride + c * col_stride]);
    uint64_t n = nrows, offset = 0, limit = 1;
    for (uint64_t i = 0; i < llv; i++) limit *= arity;
    while (llv == 0 ? n > 1 : n > limit) {
        uint64_t cur = idx % arity, si = idx - cur;
        for (uint64_t i = 0; i < arity; i++) {
            if (i == cur) continue;
            for (int k = 0; k < 4; k++) proof[w++] = tree[(offset + si + i) * 4 + k];
        }
        uint64_t next = (n + arity - 1) / arity;
        offset += next * arity;
        idx /= arity;
        n = next;
    }
    return w;
}

/* ------------------------------------------------------------------- FRI -- */
/* fri.hpp:32-99 (fold) with polMulAxi (:178-186) and evalPol (:189-206).
 * pol holds 2^polBits cubic elements (interleaved [c0,c1,c2]); out receives the
 * 2^currentBits folded elements. For every g: gather the nX = 2^(polBits-cur)
 * values pol[i*2^cur + g], interpolate (INTT of size nX), scale coefficient i by
 * sinv_g^i with sinv_g = (shift^(2^(nBitsExt-prevBits)) * w_{2^prevBits}^g)^-1, and
 * evaluate at the challenge. Step 0 is a no-op in the reference (returns 0 here). */
int oracle_fri_fold(uint64_t *out, const uint64_t *pol, const uint64_t *challenge, uint64_t step,
                    uint64_t nBitsExt, uint64_t prevBits, uint64_t currentBits)
{
    if (step == 0) return 0;
    uint64_t polBits = prevBits;
    uint64_t shift = 7;
    for (uint64_t j = 0; j < nBitsExt - prevBits; j++) shift = o_mul(shift, shift);
    uint64_t shiftInv = o_inv(shift);
    uint64_t n2 = 1ULL << currentBits, nX = (1ULL << polBits) / n2;
    uint64_t wi = o_inv(O_W[polBits]);
    uint32_t lX = o_log2(nX);
    uint64_t winvX = o_inv(O_W[lX]), nXinv = o_inv(nX);
    #pragma omp parallel for schedule(static)
    for (uint64_t g = 0; g < n2; g++) {
        /* any fold ratio: per-group buffers of nX cubic values (buf[k] = bufs + k*nX) */
        uint64_t *bufs = (uint64_t *)malloc(6 * nX * sizeof(uint64_t));
        uint64_t *buf[3] = {bufs, bufs + nX, bufs + 2 * nX};
        uint64_t (*coef)[3] = (uint64_t (*)[3])(bufs + 3 * nX);
        for (uint64_t i = 0; i < nX; i++)
            for (int k = 0; k < 3; k++) buf[k][i] = pol[(i * n2 + g) * 3 + k];
        for (int k = 0; k < 3; k++) {
            o_dft_inplace(buf[k], lX, winvX);
            for (uint64_t i = 0; i < nX; i++) coef[i][k] = o_mul(buf[k][i], nXinv);
        }
        uint64_t sinv = o_mul(shiftInv, o_pow(wi, g)), r = 1;
        for (uint64_t i = 0; i < nX; i++) {
            for (int k = 0; k < 3; k++) coef[i][k] = o_mul(coef[i][k], r);
            r = o_mul(r, sinv);
        }
        uint64_t acc[3] = {coef[nX - 1][0], coef[nX - 1][1], coef[nX - 1][2]};
        for (int64_t i = (int64_t)nX - 2; i >= 0; i--) {
            uint64_t t[3];
            o_c3mul(t, acc, challenge);
            for (int k = 0; k < 3; k++) acc[k] = o_add(t[k], coef[i][k]);
        }
        for (int k = 0; k < 3; k++) out[g * 3 + k] = acc[k];
        free(bufs);
    }
    return 0;
}

/* fri.hpp:209-226 (getTransposed): aux[(i*h + j)] = pol[(j*w + i)], w = 2^tBits,
 * h = degree / w, element = 3 u64. */
void oracle_fri_transpose(uint64_t *aux, const uint64_t *pol, uint64_t degree, uint64_t tBits)
{
    uint64_t w = 1ULL << tBits, h = degree / w;
    for (uint64_t i = 0; i < w; i++)
        for (uint64_t j = 0; j < h; j++)
            memcpy(&aux[(i * h + j) * 3], &pol[(j * w + i) * 3], 24);
}

/* Brute-force single-index fold (pins oracle_fri_fold): Lagrange-free route via
 * the coset points x_i = s * w_prev^(g + i*2^cur), solving for the coefficients of
 * the degree-<nX interpolant in the variable y = x / (s*w_prev^g) by naive inverse
 * DFT, then evaluating at challenge. Mirrors verify_fold (fri.hpp:229-256) which
 * evaluates at challenge * sinv instead of scaling coefficients. */
void oracle_fri_fold_one(uint64_t *out3, const uint64_t *vals, const uint64_t *challenge,
                         uint64_t nBitsExt, uint64_t prevBits, uint64_t currentBits, uint64_t g)
{
    uint64_t nX = 1ULL << (prevBits - currentBits);
    uint64_t shift = 7;
    for (uint64_t j = 0; j < nBitsExt - prevBits; j++) shift = o_mul(shift, shift);
    uint64_t sinv = o_inv(o_mul(shift, o_pow(O_W[prevBits], g)));
    uint64_t *mem = (uint64_t *)malloc(5 * nX * sizeof(uint64_t));
    uint64_t (*coef)[3] = (uint64_t (*)[3])mem, *col = mem + 3 * nX, *res = mem + 4 * nX;
    for (int k = 0; k < 3; k++) {
        for (uint64_t i = 0; i < nX; i++) col[i] = vals[i * 3 + k];
        oracle_dft_naive(res, col, nX, 1);
        for (uint64_t i = 0; i < nX; i++) coef[i][k] = res[i];
    }
    uint64_t x[3];
    for (int k = 0; k < 3; k++) x[k] = o_mul(challenge[k], sinv);
    uint64_t acc[3] = {coef[nX - 1][0], coef[nX - 1][1], coef[nX - 1][2]};
    for (int64_t i = (int64_t)nX - 2; i >= 0; i--) {
        uint64_t t[3];
        o_c3mul(t, acc, x);
        for (int k = 0; k < 3; k++) acc[k] = o_add(t[k], coef[i][k]);
    }
    memcpy(out3, acc, 24);
    free(mem);
}

/* --------------------------------------------------------------- helpers -- */
/* Seeded uniform elements of [0, p): counter-based SplitMix64. Element i is
 * mix(seed + (i+1)*0x9E3779B97F4A7C15); a value >= p (probability 2^-32) is
 * re-mixed until it falls below p. The same definition is used by the GPU fill
 * kernel and the numpy generator, so every layer can build the same inputs. */
static inline uint64_t o_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
void oracle_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed)
{
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t z = o_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
        while (z >= GLP) z = o_mix64(z ^ 0xD1B54A32D192ED03ULL);
        out[i] = z;
    }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
