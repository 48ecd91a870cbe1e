"""Transforms above 2^24 points, up to the reference GPU engine's 2^28
(ntt_goldilocks.cu:632). There the strided passes take two launches each (the
index bits of n1 split between them) and the twiddles between passes come from
the 2^28-domain tables. tests/golden/large_domain_reference.json holds the
compiled reference's outputs on the same seeded inputs
(tests/golden/make_fixtures.py --large): LDE digests and arity-3 roots
(NTT_Goldilocks::LDE + Poseidon2Goldilocks<12>::merkletree), plain NTT / INTT
digests, and a computeQ digest. A 2^25 LDE is also checked column by column
against the oracle."""
import json
import os

import numpy as np
import pytest

import datagen
from golden_helpers import device_colmajor_sha256

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large_domain_reference.json")


@pytest.fixture(scope="module")
def pc():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    import pil2_commit
    pil2_commit.set_hash_family(2)
    pil2_commit.set_abort_on_error(False)
    return pil2_commit


@pytest.fixture(scope="module")
def ref():
    with open(FIXTURE) as f:
        return json.load(f)


def _release():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _case(ref, key, **want):
    for c in ref[key]:
        if all(c[k] == v for k, v in want.items()):
            return c
    raise KeyError(want)


@pytest.mark.parametrize("nb,nbe,c", [(24, 25, 2), (23, 25, 3), (21, 25, 4), (27, 28, 1), (26, 28, 1)])
def test_large_lde_commit_equals_reference(pc, ref, nb, nbe, c):
    """The whole row-major LDE and the arity-3 root of a commit above 2^24 rows
    equal the compiled reference's."""
    f = _case(ref, "lde", n_bits=nb, n_bits_ext=nbe, n_cols=c)
    n, ne = 1 << nb, 1 << nbe
    print("large LDE 2^%d x %d -> 2^%d" % (nb, c, nbe), flush=True)
    trace = pc.empty(n * c)
    pc.fill_uniform(trace, f["seed"])
    lde, nodes = pc.commit(trace, 3, nb, nbe, c)
    del trace
    root = pc.MerkleTreeGL(3, 0, True, ne, c, nodes=nodes).get_root()
    assert [int(v) for v in root] == [int(x, 16) for x in f["root"]]
    assert device_colmajor_sha256(pc, lde, ne, c) == f["lde_sha256"]
    del lde, nodes
    _release()


def test_lde_2p25_columns_match_oracle(pc, oracle):
    """2^23 x 2 -> 2^25 (blowup 4: a two-launch strided pass over 11 index
    bits) against the oracle's LDE of the same columns."""
    nb, nbe, c = 23, 25, 2
    trace = datagen.uniform((1 << nb) * c, 0x2525).reshape(1 << nb, c)
    out = pc.empty(c << nbe)
    pc.NTT_Goldilocks(1 << nbe).LDE(out, pc.to_device(trace), 1 << nbe, 1 << nb, c)
    got = pc.to_host(out).reshape(c, 1 << nbe).T
    assert (got == oracle.lde(trace, 1 << nbe)).all()
    _release()


@pytest.mark.parametrize("lg", [25, 28])
def test_large_ntt_equals_reference(pc, ref, lg):
    """Plain NTT and INTT (column-major, out of place) above 2^24 points equal
    the compiled reference's NTT_Goldilocks::NTT / ::INTT."""
    f = _case(ref, "ntt", log_size=lg)
    n = 1 << lg
    print("large NTT 2^%d" % lg, flush=True)
    src = pc.empty(n)
    pc.fill_uniform(src, f["seed"])
    dst = pc.empty(n)
    ntt = pc.NTT_Goldilocks(n)
    ntt.NTT(dst, src, n, 1, row_major=False)
    assert device_colmajor_sha256(pc, dst, n, 1) == f["ntt_sha256"]
    ntt.INTT(dst, src, n, 1, row_major=False)
    assert device_colmajor_sha256(pc, dst, n, 1) == f["intt_sha256"]
    del src, dst
    _release()


def test_large_compute_q_equals_reference(pc, ref):
    """computeQ with NExt = 2^25 (q row-major, 2 pieces of 2^23) equals the
    reference's INTT / regroup / NTT."""
    f = ref["compute_q"][0]
    lb, lbe, qdeg, qdim = f["n_bits"], f["n_bits_ext"], f["q_deg"], f["q_dim"]
    ne = 1 << lbe
    q = pc.empty(ne * qdim)
    pc.fill_uniform(q, f["seed"])
    cmq, _ = pc.compute_q(q, lb, lbe, qdeg, q_dim=qdim, row_major=True, merkelize=False)
    assert device_colmajor_sha256(pc, cmq, ne, qdeg * qdim) == f["cmq_sha256"]
    del q, cmq
    _release()


def test_domain_limit(pc):
    """nBitsExt above 28 is refused with an error, not computed."""
    t = pc.empty(16)
    with pytest.raises(pc.HipError):
        pc.commit(t, 3, 4, 29, 1, lde=pc.empty(16), nodes=pc.empty(64), scratch=pc.empty(16))
