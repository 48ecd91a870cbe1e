"""Generate the golden fixtures under tests/golden/ (run in the build container only).

Two files are written:

reference_vectors.json
    Known-answer vectors held in the reference's own unit tests, read from their
    text: Poseidon2 permute goldens W8/12/16 and linearHash goldens
    (pil2-stark/src/goldilocks/tests/test_helpers.hpp:22-63) and the LDE golden of
    TEST(extendePol) (pil2-stark/src/goldilocks/tests/test_ntt_cpu.cpp:351-412).

compiled_reference.json
    Outputs of the reference's own CPU code (NTT_Goldilocks::LDE and
    Poseidon2Goldilocks<W>::merkletree, compiled by oracle/Makefile into
    oracle/_ref) on seeded inputs from datagen.uniform: Merkle roots and LDE
    digests at the plain-small shape and a sweep of shapes covering the level
    padding paths. Digest = sha256 of the canonical row-major LDE (little endian).

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
    out["_inputs"] = {
        "permute": "state = [0, 1, ..., W-1]",
        "linear_hash": "input = [1, 2, ..., 20], size 20",
        "merkletree_reduce": "3 digests [1..12], arity 3, W=12",
        "extendePol": "16 rows x 8 cols, a[0][j] = a[1][j] = 1 + j, a[i] = a[i-1] + a[i-2]; LDE to 32 rows",
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
            "cases": cases}


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
