"""Generate the golden fixtures under tests/golden/ (run in the build container only).

Two files are written:

reference_vectors.json
    Known-answer vectors held in the reference's own unit tests, read from their
    text: Poseidon2 permute goldens W8/12/16 and linearHash goldens
    (pil2-stark/src/goldilocks/tests/test_helpers.hpp:22-63), the LDE golden of
    TEST(extendePol) (pil2-stark/src/goldilocks/tests/test_ntt_cpu.cpp:351-412) and
    the INTT / shift / NTT golden of TEST(LDE) (test_ntt_cpu.cpp:182-260).

compiled_reference.json
    Outputs of the reference's own CPU code (NTT_Goldilocks::LDE, ::NTT, ::INTT and
    Poseidon2Goldilocks<W>::merkletree, compiled by oracle/Makefile into
    oracle/_ref) on seeded inputs from datagen.uniform: Merkle roots and LDE
    digests at the plain-small shape and a sweep of shapes covering the level
    padding paths; NTT / INTT digests; computeQ digests (the reference's INTT and
    NTT around the regrouping of starks.hpp:209-221). Digest = sha256 of the
    canonical row-major array (little endian).

Usage: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pil2-proofman_amd"))
from oracle.oracle import RefLib, build  # noqa: E402
import datagen  # noqa: E402

REF_TESTS = "/root/reference/pil2-stark/src/goldilocks/tests"


def _hexes(text):
    return [int(h, 16) for h in re.findall(r"0[xX][0-9a-fA-F]+", text)]


def reference_vectors():
    helpers = open(os.path.join(REF_TESTS, "test_helpers.hpp")).read()
    out = {}
    for w in (8, 12, 16):
        m = re.search(r"PERMUTE_W%d_GOLDEN\[%d\]\s*=\s*\{(.*?)\};" % (w, w), helpers, re.S)
        out["permute_w%d" % w] = ["0x%016x" % v for v in _hexes(m.group(1))]
        m = re.search(r"LINEAR_HASH_W%d_GOLDEN\[4\]\s*=\s*\{(.*?)\};" % w, helpers, re.S)
        out["linear_hash_w%d_1to20" % w] = ["0x%016x" % v for v in _hexes(m.group(1))]
    m = re.search(r"MERKLETREE_REDUCE_W12_AR3_GOLDEN\[4\]\s*=\s*\{(.*?)\};", helpers, re.S)
    out["merkletree_reduce_w12_ar3_1to12"] = ["0x%016x" % v for v in _hexes(m.group(1))]
    ntt = open(os.path.join(REF_TESTS, "test_ntt_cpu.cpp")).read()
    body = ntt[ntt.index("TEST(GOLDILOCKS_TEST, extendePol)"):ntt.index("TEST(GOLDILOCKS_TEST, LDE_correctness)")]
    vals = re.findall(r"ASSERT_EQ\(Goldilocks::toU64\(a\[(\d+) \* NUM_COLUMNS\]\), (0X[0-9A-F]+)\)", body)
    out["extendePol_16x8_to_32_col0"] = ["0x%016x" % int(v, 16) for _, v in sorted(vals, key=lambda t: int(t[0]))]
    body = ntt[ntt.index("TEST(GOLDILOCKS_TEST, LDE)\n"):ntt.index("TEST(GOLDILOCKS_TEST, LDE_block)")]
    vals = re.findall(r"ASSERT_EQ\(Goldilocks::toU64\(a\[(\d+)\]\), (0X[0-9A-F]+)\)", body)
    out["intt_shift49_ntt_fib16_to_32"] = ["0x%016x" % int(v, 16) for _, v in sorted(vals, key=lambda t: int(t[0]))]
    out["_inputs"] = {
        "permute": "state = [0, 1, ..., W-1]",
        "linear_hash": "input = [1, 2, ..., 20], size 20",
        "merkletree_reduce": "3 digests [1..12], arity 3, W=12",
        "extendePol": "16 rows x 8 cols, a[0][j] = a[1][j] = 1 + j, a[i] = a[i-1] + a[i-2]; LDE to 32 rows",
        "intt_shift49": "a[0] = a[1] = 1, a[i] = a[i-1] + a[i-2] (16 points); INTT(16), a[i] *= 49^i, "
                        "zero-extend to 32, NTT(32) (TEST(GOLDILOCKS_TEST, LDE), test_ntt_cpu.cpp:182-260)",
    }
    return out


def digest(lde_rowmajor):
    return hashlib.sha256(np.ascontiguousarray(lde_rowmajor, dtype="<u8").tobytes()).hexdigest()


def compiled_reference(ref):
    cases = []
    # plain small: 2^16 x 32, blowup 2 at all three arities
    shapes = [(16, 17, 32, 0x5EED0001, (2, 3, 4))]
    # sweep: level padding / extraZeros paths, odd column counts, several blowups
    for lb, lbe, c in [(12, 13, 8), (12, 14, 13), (13, 14, 32), (14, 15, 256), (10, 12, 1), (11, 13, 3),
                       (9, 10, 100), (12, 12, 7), (15, 17, 9), (3, 5, 5), (0, 1, 4), (6, 8, 12)]:
        shapes.append((lb, lbe, c, 0xA000 + lb * 64 + lbe * 8 + c, (2, 3, 4)))
    for lb, lbe, c, seed, arities in shapes:
        n, ne = 1 << lb, 1 << lbe
        trace = datagen.uniform(n * c, seed).reshape(n, c)
        lde = ref.lde(trace, ne)
        case = {"n_bits": lb, "n_bits_ext": lbe, "n_cols": c, "seed": seed, "lde_sha256": digest(lde),
                "lde_first_row": ["0x%016x" % v for v in lde[0]], "lde_last_row": ["0x%016x" % v for v in lde[-1]],
                "roots": {}}
        for a in arities:
            from oracle.oracle import Oracle
            tl = Oracle().tree_num_elements(ne, a)
            tree = ref.merkletree(lde, a, tl)
            case["roots"][str(a)] = ["0x%016x" % v for v in tree[-4:]]
        cases.append(case)
        print("case", lb, lbe, c, "done", flush=True)
    return {"generator": "datagen.uniform(n*cols, seed) row-major", "library": os.path.basename(ref.path),
            "cases": cases, "ntt": ntt_cases(ref), "compute_q": compute_q_cases(ref)}


def ntt_cases(ref):
    """NTT_Goldilocks::NTT and ::INTT on seeded row-major size x ncols arrays."""
    out = []
    for lg, c in [(0, 3), (1, 3), (4, 8), (10, 5), (13, 2), (16, 1)]:
        seed = 0xB000 + lg * 16 + c
        a = datagen.uniform((1 << lg) * c, seed).reshape(1 << lg, c)
        out.append({"log_size": lg, "n_cols": c, "seed": seed, "ntt_sha256": digest(ref.ntt(a)),
                    "intt_sha256": digest(ref.ntt(a, inverse=True))})
    return out


def compute_q_cases(ref):
    """Starks::computeQ (starks.hpp:187-239) driven through the reference's own
    NTT_Goldilocks::INTT / NTT: q (row-major NExt x qDim, seeded) -> INTT over NExt,
    cmQ[i][p*qDim + d] = coef[p*N + i][d] * (1/7)^(N*p) for i < N, zero rows
    N..NExt, NTT over NExt. Digest of the row-major cmQ."""
    P = 0xFFFFFFFF00000001
    out = []
    for lb, lbe, qdeg, qdim in [(0, 2, 2, 3), (2, 3, 2, 3), (3, 5, 3, 3), (4, 5, 2, 3), (10, 12, 4, 3),
                                (12, 13, 2, 3), (14, 16, 3, 3), (11, 13, 1, 1), (16, 17, 2, 3)]:
        n, ne = 1 << lb, 1 << lbe
        seed = 0xC000 + lb * 64 + lbe * 8 + qdeg
        q = datagen.uniform(ne * qdim, seed).reshape(ne, qdim)
        coef = ref.ntt(q, inverse=True)
        shift_in = pow(pow(7, P - 2, P), n, P)
        cmq = np.zeros((ne, qdeg * qdim), dtype=np.uint64)
        for p in range(qdeg):
            s = pow(shift_in, p, P)
            for d in range(qdim):
                col = [(int(v) * s) % P for v in coef[p * n:(p + 1) * n, d]]
                cmq[:n, p * qdim + d] = np.array(col, dtype=np.uint64)
        cmq = ref.ntt(cmq)
        out.append({"n_bits": lb, "n_bits_ext": lbe, "q_deg": qdeg, "q_dim": qdim, "seed": seed,
                    "cmq_sha256": digest(cmq), "cmq_row1": ["0x%016x" % v for v in cmq[1]]})
        print("compute_q case", lb, lbe, qdeg, qdim, "done", flush=True)
    return out


def main():
    build()
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(reference_vectors(), f, indent=1)
    ref = RefLib()
    with open(os.path.join(HERE, "compiled_reference.json"), "w") as f:
        json.dump(compiled_reference(ref), f, indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main()
