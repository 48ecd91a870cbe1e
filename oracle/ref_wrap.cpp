// C-ABI wrapper over the reference's own CPU classes (TEST INFRASTRUCTURE ONLY).
//
// Compiled by oracle/Makefile directly against the reference's sources; the
// reference sources are never copied into this repository. The resulting
// oracle/_ref/libref_*.so is used by tests/ to pin the C restatement
// (oracle/gl_oracle.c) and by bench.py as the timed CPU baseline.
//
// Each wrapper names the reference entry point it calls:
//   ref_lde        -> NTT_Goldilocks(N).LDE          ntt_goldilocks.cpp:369-404
//   ref_ntt        -> NTT_Goldilocks(n).NTT / INTT   ntt_goldilocks.cpp:211-260, ntt_goldilocks.hpp:187-191
//   ref_compute_q  -> Starks::computeQ's transform steps (starks.hpp:196-226) on
//                     NTT_Goldilocks::INTT / NTT and Goldilocks::mul
//   ref_merkletree -> Poseidon2Goldilocks<4a>::merkletree (Auto mode, as
//                     MerkleTreeGL::merkelize does)  merkleTreeGL.cpp:316-336
//   ref_permute    -> Poseidon2Goldilocks<W>::permute poseidon2_goldilocks.hpp:252-272
//   ref_linear_hash-> Poseidon2Goldilocks<W>::linearHash
//   ref_grinding   -> Poseidon2GoldilocksGrinding::grinding (width 8)
//   ref_c3_mul     -> Goldilocks3::mul               goldilocks_cubic_extension.hpp:232-245
#include "goldilocks_base_field.hpp"
#include "goldilocks_cubic_extension.hpp"
#include "ntt_goldilocks.hpp"
#include "poseidon2_goldilocks.hpp"
#include <omp.h>
#include <vector>

template <uint32_t W>
static void perm(uint64_t *st)
{
    Goldilocks::Element in[W], out[W];
    for (uint32_t i = 0; i < W; i++) in[i].fe = st[i];
    Poseidon2Goldilocks<W>::permute(out, in, Poseidon2Mode::Scalar);
    for (uint32_t i = 0; i < W; i++) st[i] = Goldilocks::toU64(out[i]);
}

extern "C" {

void ref_lde(uint64_t *out, const uint64_t *in, uint64_t NExt, uint64_t N, uint64_t ncols)
{
    NTT_Goldilocks ntt(N);
    ntt.LDE((Goldilocks::Element *)out, (Goldilocks::Element *)in, NExt, N, ncols);
}

// in place on a row-major size x ncols array
void ref_ntt(uint64_t *a, uint64_t size, uint64_t ncols, int inverse)
{
    NTT_Goldilocks ntt(size);
    Goldilocks::Element *x = (Goldilocks::Element *)a;
    if (inverse)
        ntt.INTT(x, x, size, ncols);
    else
        ntt.NTT(x, x, size, ncols);
}

// q: row-major NExt x qDim (transformed in place, as in the reference);
// cmq: row-major NExt x qDeg*qDim
void ref_compute_q(uint64_t *cmq, uint64_t *q, uint64_t N, uint64_t NExt, uint64_t qDeg, uint64_t qDim)
{
    NTT_Goldilocks ntt(NExt);
    Goldilocks::Element *qe = (Goldilocks::Element *)q, *cm = (Goldilocks::Element *)cmq;
    ntt.INTT(qe, qe, NExt, qDim);
    Goldilocks::Element shiftIn = Goldilocks::exp(Goldilocks::inv(Goldilocks::shift()), N);
    std::vector<Goldilocks::Element> S(qDeg);
    S[0] = Goldilocks::one();
    for (uint64_t p = 1; p < qDeg; p++) S[p] = Goldilocks::mul(S[p - 1], shiftIn);
    uint64_t nc = qDeg * qDim;
#pragma omp parallel for collapse(2)
    for (uint64_t p = 0; p < qDeg; p++)
        for (uint64_t i = 0; i < N; i++)
            for (uint64_t d = 0; d < qDim; d++) cm[i * nc + p * qDim + d] = Goldilocks::mul(qe[(p * N + i) * qDim + d], S[p]);
#pragma omp parallel for
    for (uint64_t k = N * nc; k < NExt * nc; k++) cm[k] = Goldilocks::zero();
    ntt.NTT(cm, cm, NExt, nc);
}

int ref_merkletree(uint64_t *tree, const uint64_t *src, uint64_t ncols, uint64_t nrows, uint64_t arity)
{
    Goldilocks::Element *t = (Goldilocks::Element *)tree;
    Goldilocks::Element *s = (Goldilocks::Element *)src;
    switch (arity) {
    case 2: Poseidon2Goldilocks<8>::merkletree(t, s, ncols, nrows, arity); return 0;
    case 3: Poseidon2Goldilocks<12>::merkletree(t, s, ncols, nrows, arity); return 0;
    case 4: Poseidon2Goldilocks<16>::merkletree(t, s, ncols, nrows, arity); return 0;
    default: return -1;
    }
}

int ref_permute(uint64_t *st, uint32_t W)
{
    switch (W) {
    case 8: perm<8>(st); return 0;
    case 12: perm<12>(st); return 0;
    case 16: perm<16>(st); return 0;
    default: return -1;
    }
}

int ref_linear_hash(uint64_t *out4, const uint64_t *in, uint64_t size, uint32_t W)
{
    Goldilocks::Element o[4];
    Goldilocks::Element *i = (Goldilocks::Element *)in;
    switch (W) {
    case 8: Poseidon2Goldilocks<8>::linearHash(o, i, size, Poseidon2Mode::Scalar); break;
    case 12: Poseidon2Goldilocks<12>::linearHash(o, i, size, Poseidon2Mode::Scalar); break;
    case 16: Poseidon2Goldilocks<16>::linearHash(o, i, size, Poseidon2Mode::Scalar); break;
    default: return -1;
    }
    for (int k = 0; k < 4; k++) out4[k] = Goldilocks::toU64(o[k]);
    return 0;
}

// Poseidon2Goldilocks<8>::grinding (poseidon2_goldilocks.cpp:177-232); UINT64_MAX on failure
uint64_t ref_grinding(const uint64_t *ch, uint32_t n_bits)
{
    uint64_t nonce = UINT64_MAX;
    try {
        Poseidon2GoldilocksGrinding::grinding(nonce, ch, n_bits);
    } catch (...) {
        return UINT64_MAX;
    }
    return nonce;
}

void ref_c3_mul(uint64_t *r, const uint64_t *a, const uint64_t *b)
{
    Goldilocks3::Element x, y, z;
    for (int k = 0; k < 3; k++) { x[k].fe = a[k]; y[k].fe = b[k]; }
    Goldilocks3::mul(z, x, y);
    for (int k = 0; k < 3; k++) r[k] = Goldilocks::toU64(z[k]);
}

int ref_num_threads(void) { return omp_get_max_threads(); }

}
