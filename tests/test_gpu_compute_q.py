"""GPU parity tests for the quotient commit (pil2hip_compute_q): Starks::computeQ
(starks.hpp:187-239) plus the step's merkelize, through the C-ABI, against the
oracle, the compiled reference's fixtures and, at the full 2^22 -> 2^24 shape,
the quotient-split identity. Integer work: every comparison is exact."""
import numpy as np
import pytest

from conftest import hexes
from golden_helpers import colmajor_to_rowmajor, lde_digest
import datagen

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def pc():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    import pil2_commit
    pil2_commit.set_hash_family(2)
    pil2_commit.set_abort_on_error(False)
    return pil2_commit


def _sync():
    import torch
    torch.cuda.synchronize()


def gpu_compute_q(pc, q, nb, nbe, qdeg, arity=3, row_major=True):
    """q: host row-major (NExt, qDim). Returns (cmQ row-major host, nodes host)."""
    ne, qdim = q.shape
    src = q if row_major else np.ascontiguousarray(q.T)
    dq = pc.to_device(src)
    cm, nodes = pc.compute_q(dq, nb, nbe, qdeg, q_dim=qdim, arity=arity, row_major=row_major)
    _sync()
    assert (pc.to_host(dq) == src).all()  # the input is left untouched
    return colmajor_to_rowmajor(pc.to_host(cm), ne, qdeg * qdim), pc.to_host(nodes)


def test_compute_q_matches_compiled_reference(pc, compiled_ref, oracle):
    """Every computeQ fixture (the reference's INTT / NTT around the regrouping):
    N = 1 and NExt < 16 (direct path), single-pass and strided plans, qDim 1."""
    for case in compiled_ref["compute_q"]:
        nb, nbe, qdeg, qdim = case["n_bits"], case["n_bits_ext"], case["q_deg"], case["q_dim"]
        print("computeQ case", nb, nbe, qdeg, qdim, flush=True)
        q = datagen.uniform((1 << nbe) * qdim, case["seed"]).reshape(1 << nbe, qdim)
        cm, nodes = gpu_compute_q(pc, q, nb, nbe, qdeg)
        assert [int(v) for v in cm[1]] == hexes(case["cmq_row1"]), case
        assert lde_digest(cm) == case["cmq_sha256"], case
        if (1 << nbe) * qdeg * qdim <= (1 << 16):
            assert (nodes == oracle.merkletree(cm, 3)).all(), case


@pytest.mark.parametrize("nb,nbe,qdeg,qdim,arity", [(17, 19, 4, 3, 3), (13, 14, 1, 3, 2), (9, 13, 16, 3, 4),
                                                     (5, 9, 3, 2, 3), (20, 21, 2, 3, 3)])
def test_compute_q_matches_oracle(pc, oracle, nb, nbe, qdeg, qdim, arity):
    """Plans with a strided pass (n1 > 1), blowups 2..16, qDeg below and at the
    blowup; both input layouts give the same cmQ; the root is the oracle's."""
    ne = 1 << nbe
    q = datagen.uniform(ne * qdim, 0xE000 + nb * 64 + nbe).reshape(ne, qdim)
    want = oracle.compute_q(q, 1 << nb, qdeg)
    cm, nodes = gpu_compute_q(pc, q, nb, nbe, qdeg, arity=arity)
    assert (cm == want).all()
    cm2, nodes2 = gpu_compute_q(pc, q, nb, nbe, qdeg, arity=arity, row_major=False)
    assert (cm2 == want).all()
    assert (nodes2 == nodes).all()
    if ne * qdeg * qdim <= (1 << 20):
        assert (nodes[-4:] == oracle.merkletree(want, arity)[-4:]).all()


def test_compute_q_full_shape_split_identity(pc, oracle):
    """2^22 -> 2^24, qDeg 4, qDim 3. q is the coset evaluation of a random
    polynomial of degree < 2^24 (an LDE from 2^24 rows to 2^24), so the output
    columns are the unique split q(x) = sum_p x^(p*N) Q_p(x): checked at 4096
    sampled rows x_j = 7 w^j; the tree root matches a re-merkelization."""
    import torch
    nb, nbe, qdeg, qdim = 22, 24, 4, 3
    n, ne = 1 << nb, 1 << nbe
    trace = pc.empty(ne * qdim)
    pc.fill_uniform(trace, 0x5EED0006)
    q = pc.empty(ne * qdim)
    pc.NTT_Goldilocks(ne).LDE(q, trace, ne, ne, qdim)   # column-major q
    del trace
    cm, nodes = pc.compute_q(q, nb, nbe, qdeg, q_dim=qdim, arity=3, row_major=False)
    _sync()
    rows = np.random.default_rng(5).choice(ne, 4096, replace=False)
    rows_t = torch.as_tensor(rows.astype(np.int64), device=q.device)
    qs = pc.to_host(q.view(qdim, ne)[:, rows_t].contiguous()).reshape(qdim, -1)
    cs = pc.to_host(cm.view(qdeg * qdim, ne)[:, rows_t].contiguous()).reshape(qdeg * qdim, -1)
    w = oracle.root(nbe)
    for k, j in enumerate(rows):
        xn = pow((7 * pow(w, int(j), P)) % P, n, P)
        for d in range(qdim):
            acc = sum(int(cs[p * qdim + d, k]) * pow(xn, p, P) for p in range(qdeg)) % P
            assert acc == int(qs[d, k]), (int(j), d)
    root = pc.to_host(nodes)[-4:].copy()
    nodes2 = pc.empty(nodes.numel())
    assert pc.lib.pil2hip_merkelize(pc._ptr(nodes2), pc._ptr(cm), ne, qdeg * qdim, 1, ne, 3, pc._stream()) == 0
    _sync()
    assert (pc.to_host(nodes2)[-4:] == root).all()


def test_compute_q_errors(pc):
    q = pc.empty(64 * 3)
    cm = pc.empty(64 * 12)
    sc = pc.empty(64 * 12)
    L = pc.lib
    # qDeg above the blowup
    assert L.pil2hip_compute_q(pc._ptr(cm), None, pc._ptr(q), 3, 1, pc._ptr(sc), 4, 6, 5, 3, 3, pc._stream()) < 0
    # neither row- nor column-major
    assert L.pil2hip_compute_q(pc._ptr(cm), None, pc._ptr(q), 3, 2, pc._ptr(sc), 4, 6, 4, 3, 3, pc._stream()) < 0
    # column-major with a column stride shorter than NExt
    assert L.pil2hip_compute_q(pc._ptr(cm), None, pc._ptr(q), 1, 32, pc._ptr(sc), 4, 6, 4, 3, 3, pc._stream()) < 0
    assert L.pil2hip_compute_q(pc._ptr(cm), None, pc._ptr(q), 3, 1, pc._ptr(sc), 4, 6, 0, 3, 3, pc._stream()) < 0
    assert "qDeg" in pc.lib.pil2hip_last_error().decode()
