"""GPU parity tests: libpil2_commit_hip.so (through its C-ABI) against the oracle
and the fixtures produced by the compiled reference. Integer work, so every
comparison is exact (canonical values)."""
import os

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


def gpu_lde(pc, trace, nb, nbe):
    n, c = trace.shape
    out = pc.empty((1 << nbe) * c)
    pc.NTT_Goldilocks(n).LDE(out, pc.to_device(trace), 1 << nbe, n, c)
    _sync()
    return colmajor_to_rowmajor(pc.to_host(out), 1 << nbe, c)


# ----------------------------------------------------------------- field --
EDGE = [0, 1, 2, 3, 7, 2**31, 2**32 - 2, 2**32 - 1, 2**32, 2**32 + 1, 2**63 - 1, 2**63, 2**63 + 1,
        P - 2**32, P - 2, P - 1]
RAW_EDGE = EDGE + [P, P + 1, P + 2**31, 2**64 - 2, 2**64 - 1]  # non-canonical words (mul accepts any)


def _binop(pc, op, a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    out = pc.empty(a.size)
    da, db = pc.to_device(a), pc.to_device(b)
    rc = pc.lib.pil2hip_gl_binop(pc._ptr(out), pc._ptr(da), pc._ptr(db), a.size, op, pc._stream())
    assert rc == 0
    _sync()
    return [int(v) for v in pc.to_host(out)]


def _pairs(vals, rng, n_random):
    a = [x for x in vals for _ in vals] + [int(v) for v in rng.integers(0, 2**64, n_random, dtype=np.uint64)]
    b = [y for _ in vals for y in vals] + [int(v) for v in rng.integers(0, 2**64, n_random, dtype=np.uint64)]
    return a, b


def test_field_ops_edge_values(pc):
    """Device add/sub/mul/reduce against exact integer arithmetic on all pairs of
    edge words (carries, borrows, values near p and 2^64) plus random words."""
    rng = np.random.default_rng(11)
    a, b = _pairs(RAW_EDGE, rng, 20000)
    assert _binop(pc, 2, a, b) == [(x * y) % P for x, y in zip(a, b)]
    assert _binop(pc, 0, a, b) == [(x + y) % P for x, y in zip(a, b)]
    assert _binop(pc, 1, a, b) == [(x - y) % P for x, y in zip(a, b)]
    # 128-bit reduction with words chosen to hit every carry/borrow combination
    words = [0, 1, 2**32 - 1, 2**32, 2**63, 2**64 - 2**32, 2**64 - 2**32 + 1, 2**64 - 1]
    lo, hi = _pairs(words, rng, 20000)
    assert _binop(pc, 3, lo, hi) == [((h << 64) + l) % P for l, h in zip(lo, hi)]


def test_mul_peak_closed_form(pc):
    """The product calibration kernel (the Poseidon2 kernels' own reduced product,
    chained) against its closed form a * b^iters mod p, over edge words (non-canonical
    inputs included) and random words, on a lane count that leaves a partial block."""
    rng = np.random.default_rng(12)
    c = pc.MUL_PEAK_CHAINS
    lanes, iters = 1000, 37
    words = [int(v) for v in rng.integers(0, 2**64, 2 * c * lanes, dtype=np.uint64)]
    for k, v in enumerate(RAW_EDGE):  # edge words as both multiplicands and multipliers
        words[2 * c * k + k % c] = v
        words[2 * c * k + c + (k + 1) % c] = v
    inp = pc.to_device(np.array(words, dtype=np.uint64))
    out = pc.empty(c * lanes)
    pc.mul_peak(out, inp, lanes, iters)
    got = pc.to_host(out).tolist()
    want = [words[2 * c * i + k] * pow(words[2 * c * i + c + k], iters, P) % P for i in range(lanes) for k in range(c)]
    assert got == want
    pc.mul_peak(out, inp, lanes, 0)  # zero steps: the canonical inputs
    assert pc.to_host(out).tolist() == [words[2 * c * i + k] % P for i in range(lanes) for k in range(c)]


# ------------------------------------------------------------- Poseidon2 --
@pytest.mark.parametrize("w", [8, 12, 16])
def test_permute_reference_goldens(pc, refvec, w):
    st = pc.to_device(np.arange(w, dtype=np.uint64))
    pc.p2_permute(st, w)
    _sync()
    assert [int(v) for v in pc.to_host(st)] == hexes(refvec["permute_w%d" % w])


def test_permute_random_batch(pc, oracle):
    for w in (8, 12, 16):
        states = datagen.uniform(w * 1000, w)
        d = pc.to_device(states)
        pc.p2_permute(d, w)
        _sync()
        got = pc.to_host(d).reshape(1000, w)
        for i in (0, 1, 500, 999):
            assert (got[i] == oracle.permute(states[i * w:(i + 1) * w])).all()


# words at the carry and borrow boundaries of the non-canonical arithmetic,
# including values >= p (any 64-bit word is a valid field representative)
EDGE_WORDS = [0, 1, 2, 2**31, 2**32 - 1, 2**32, 2**32 + 1, 2**63 - 1, 2**63, P - 2**32, P - 2, P - 1, P, P + 1,
              P + 2**32 - 2, 2**64 - 2**33, 2**64 - 2, 2**64 - 1]


def test_permute_edge_words(pc, oracle):
    """States drawn from boundary words (and all-equal states of each) through
    the permutation, every state checked against the oracle."""
    rng = np.random.default_rng(2024)
    for w in (8, 12, 16):
        n_rand = 600
        states = np.array(EDGE_WORDS, dtype=np.uint64)[rng.integers(0, len(EDGE_WORDS), (n_rand, w))]
        const = np.repeat(np.array(EDGE_WORDS, dtype=np.uint64)[:, None], w, axis=1)
        allst = np.concatenate([states, const]).astype(np.uint64)
        d = pc.to_device(allst.reshape(-1))
        pc.p2_permute(d, w)
        _sync()
        got = pc.to_host(d).reshape(-1, w)
        for i in range(allst.shape[0]):
            assert (got[i] == oracle.permute(allst[i])).all(), (w, i, [hex(int(v)) for v in allst[i]])


def test_merkle_edge_words(pc, oracle):
    """Rows of boundary words (non-canonical included) hashed into a tree."""
    rng = np.random.default_rng(7)
    rows, cols = 1 << 9, 21
    src = np.array(EDGE_WORDS, dtype=np.uint64)[rng.integers(0, len(EDGE_WORDS), (rows, cols))]
    for arity in (2, 3, 4):
        t = pc.MerkleTreeGL(arity, 0, True, rows, cols, source=pc.to_device(src))
        t.merkelize()
        _sync()
        assert (pc.to_host(t.nodes) == oracle.merkletree(src, arity)).all(), arity


# ------------------------------------------------------------------- LDE --
LDE_SHAPES = [(0, 1, 4), (1, 1, 3), (1, 3, 2), (3, 5, 5), (4, 5, 8), (6, 8, 12), (10, 12, 1), (12, 13, 8),
              (12, 12, 7), (13, 14, 32), (11, 13, 3), (14, 16, 5), (15, 16, 4), (16, 17, 32), (17, 19, 3),
              (18, 19, 2), (19, 20, 1), (20, 22, 1)]


@pytest.mark.parametrize("nb,nbe,c", LDE_SHAPES)
def test_lde_matches_oracle(pc, oracle, nb, nbe, c):
    print("lde case", nb, nbe, c, flush=True)
    trace = datagen.uniform((1 << nb) * c, 0xC0FFEE + nb * 100 + nbe * 10 + c).reshape(1 << nb, c)
    got = gpu_lde(pc, trace, nb, nbe)
    assert (got == oracle.lde(trace, 1 << nbe)).all()


def test_lde_noncanonical_input(pc, oracle):
    nb, nbe, c = 8, 9, 4
    trace = datagen.uniform((1 << nb) * c, 5).reshape(1 << nb, c)
    small = trace < np.uint64(2**32 - 1)
    raw = trace.copy()
    raw[small] += np.uint64(P)  # same residue, non-canonical word
    assert (gpu_lde(pc, raw, nb, nbe) == oracle.lde(trace, 1 << nbe)).all()


# ---------------------------------------------------------------- commit --
def test_commit_roots_match_compiled_reference(pc, compiled_ref):
    for case in compiled_ref["cases"]:
        nb, nbe, c = case["n_bits"], case["n_bits_ext"], case["n_cols"]
        print("fixture case", nb, nbe, c, flush=True)
        trace = datagen.uniform((1 << nb) * c, case["seed"]).reshape(1 << nb, c)
        d_trace = pc.to_device(trace)
        for a, root in case["roots"].items():
            lde, nodes = pc.commit(d_trace, int(a), nb, nbe, c)
            t = pc.MerkleTreeGL(int(a), 0, True, 1 << nbe, c, source=lde, row_stride=1, col_stride=1 << nbe,
                                nodes=nodes)
            assert [int(v) for v in t.get_root()] == hexes(root), (nb, nbe, c, a)
        got = colmajor_to_rowmajor(pc.to_host(lde), 1 << nbe, c)
        assert lde_digest(got) == case["lde_sha256"]
        assert [int(v) for v in got[0]] == hexes(case["lde_first_row"])


@pytest.mark.parametrize("arity", [2, 3, 4])
@pytest.mark.parametrize("rows,cols", [(1, 5), (2, 3), (3, 8), (5, 9), (7, 0), (100, 17), (1000, 12), (4096, 1),
                                       (2187, 24)])
def test_merkelize_rowmajor_full_tree(pc, oracle, arity, rows, cols):
    src = datagen.uniform(max(rows * cols, 1), rows * 31 + cols)[:rows * cols].reshape(rows, cols)
    d = pc.to_device(src if src.size else np.zeros(1, np.uint64))
    t = pc.MerkleTreeGL(arity, 0, True, rows, cols, source=d)
    t.merkelize()
    _sync()
    assert (pc.to_host(t.nodes) == oracle.merkletree(src, arity)).all()


def test_commit_is_deterministic(pc):
    nb, nbe, c = 14, 16, 24
    d = pc.empty((1 << nb) * c)
    pc.fill_uniform(d, 0x5EED0001)
    roots = []
    for _ in range(3):
        lde, nodes = pc.commit(d, 3, nb, nbe, c)
        roots.append(pc.MerkleTreeGL(3, 0, True, 1 << nbe, c, nodes=nodes).get_root())
    assert all((r == roots[0]).all() for r in roots)


@pytest.mark.parametrize("nb,nbe,c,arity,gc", [
    (10, 12, 13, 3, 0),    # automatic groups, ragged last group
    (10, 11, 40, 3, 8),    # five full groups of one rate block
    (12, 14, 40, 3, 16),   # 16 + 16 + 8
    (9, 11, 7, 2, 4),      # width 8, rate 4
    (11, 12, 33, 4, 12),   # width 16, rate 12, ragged
    (10, 12, 8, 3, 8),     # one group
    (13, 15, 24, 3, 8),    # 2^15 rows: one-lane-per-row leaf kernel
    (8, 9, 0, 3, 0),       # no columns: zero digests
    (20, 21, 64, 3, 0),    # 512 MiB of trace: automatic width, four groups of 16
])
def test_commit_host_matches_device_commit(pc, nb, nbe, c, arity, gc):
    """The streamed host commit (column groups copied on their own stream and
    absorbed with the capacity carried between groups) gives pil2hip_commit's LDE
    and tree, on repeated calls."""
    import torch
    print("commit_host case", nb, nbe, c, arity, gc, flush=True)
    n, ne = 1 << nb, 1 << nbe
    d = pc.empty(max(n * c, 1))
    pc.fill_uniform(d, 0xC0FFEE + c)
    lde_ref, nodes_ref = pc.commit(d, arity, nb, nbe, c)
    torch.cuda.synchronize()
    want_lde, want_nodes = pc.to_host(lde_ref)[:ne * c], pc.to_host(nodes_ref)
    pinned = torch.empty(max(n * c, 1), dtype=torch.int64, pin_memory=True)
    pinned.copy_(d.cpu())
    for _ in range(2):
        lde, nodes = pc.commit_host(pinned, arity, nb, nbe, c, group_cols=gc)
        torch.cuda.synchronize()
        assert np.array_equal(pc.to_host(nodes), want_nodes)
        assert np.array_equal(pc.to_host(lde)[:ne * c], want_lde)


def test_commit_host_rejects_group_width_off_the_rate(pc):
    import torch
    h = torch.zeros(1 << 10, dtype=torch.int64, pin_memory=True)
    with pytest.raises(pc.HipError, match="multiple of the sponge rate"):
        pc.commit_host(h, 3, 6, 7, 16, group_cols=5)


def test_commit_host_rejects_pageable_memory(pc):
    """A pageable host trace is refused before any copy (-34), and the next
    call on pinned memory still works (no stale error from the pointer query)."""
    import torch
    n, c = 1 << 8, 16
    pageable = torch.zeros(n * c, dtype=torch.int64)
    with pytest.raises(pc.HipError, match="page-locked"):
        pc.commit_host(pageable, 3, 8, 9, c)
    pinned = torch.zeros(n * c, dtype=torch.int64, pin_memory=True)
    lde, nodes = pc.commit_host(pinned, 3, 8, 9, c)
    torch.cuda.synchronize()
    d = pc.empty(n * c)
    d.zero_()
    lde_ref, nodes_ref = pc.commit(d, 3, 8, 9, c)
    torch.cuda.synchronize()
    assert torch.equal(nodes, nodes_ref)


def test_fill_uniform_matches_datagen(pc):
    d = pc.empty(100003)
    pc.fill_uniform(d, 0x5EED0002, start=11)
    _sync()
    assert (pc.to_host(d) == datagen.uniform(100003, 0x5EED0002, start=11)).all()


# ---------------------------------------------------- write_custom_commit --
def test_write_custom_commit_plain_small(pc, compiled_ref, oracle, tmp_path):
    case = compiled_ref["cases"][0]
    nb, nbe, c = case["n_bits"], case["n_bits_ext"], case["n_cols"]
    assert (nb, nbe, c) == (16, 17, 32)
    trace = datagen.uniform((1 << nb) * c, case["seed"]).reshape(1 << nb, c)
    for a in ("3", "4"):
        root = np.zeros(4, dtype=np.uint64)
        f = str(tmp_path / ("cc_%s.bin" % a))
        pc.write_custom_commit(root, int(a), nb, nbe, c, None, trace, f)
        assert [int(v) for v in root] == hexes(case["roots"][a])
        raw = np.fromfile(f, dtype="<u8")
        n, ne = 1 << nb, 1 << nbe
        ntree = pc.get_tree_num_elements(ne, int(a))
        assert raw.size == 4 + n * c + ne * c + ntree
        assert (raw[:4] == root).all()
        assert (raw[4:4 + n * c] == trace.ravel()).all()
        lde = raw[4 + n * c:4 + n * c + ne * c].reshape(ne, c)
        assert lde_digest(lde) == case["lde_sha256"]
        assert (raw[4 + n * c + ne * c:][-4:] == root).all()


def test_write_custom_commit_with_device_buffers_handle(pc, compiled_ref):
    case = compiled_ref["cases"][1]
    nb, nbe, c = case["n_bits"], case["n_bits_ext"], case["n_cols"]
    trace = datagen.uniform((1 << nb) * c, case["seed"]).reshape(1 << nb, c)
    h = pc.gen_device_buffers(3, nbe)
    try:
        root = np.zeros(4, dtype=np.uint64)
        pc.write_custom_commit(root, 3, nb, nbe, c, h, trace, "")
        assert [int(v) for v in root] == hexes(case["roots"]["3"])
    finally:
        pc.free_device_buffers(h)


def test_error_paths_report_instead_of_abort(pc):
    trace = np.zeros((16, 2), dtype=np.uint64)
    root = np.zeros(4, dtype=np.uint64)
    try:
        pc.set_hash_family(1)
        with pytest.raises(pc.HipError, match="Poseidon1"):
            pc.write_custom_commit(root, 3, 4, 5, 2, None, trace, "")
        pc.set_hash_family(2)
        with pytest.raises(pc.HipError, match="arity"):
            pc.write_custom_commit(root, 5, 4, 5, 2, None, trace, "")
        with pytest.raises(pc.HipError, match="nBitsExt"):
            pc.write_custom_commit(root, 3, 5, 4, 2, None, trace, "")
    finally:
        pc.set_hash_family(2)


# ------------------------------------------------------------------- FRI --
def fri_steps(nbe, folding=3, final_degree=5):
    """FRI step schedule (setup/pil2-stark/src/types/stark_struct.rs:162-166)."""
    steps = [nbe]
    b = nbe
    while b > final_degree + 1:
        b = b - folding if b > folding + final_degree else final_degree
        steps.append(b)
    return steps


@pytest.mark.parametrize("nbe", [12, 16, 22])
def test_fri_fold_all_layers(pc, oracle, nbe):
    """FRI config (seed 0x5EED0004): fold + per-layer Merkle commit, all layers and
    all roots exact vs the oracle. nbe=22 is the full FRI fold configuration."""
    steps = fri_steps(nbe)
    if nbe == 22:
        assert steps == [22, 19, 16, 13, 10, 7, 5]
    pol = datagen.uniform(3 << nbe, 0x5EED0004)
    chal = datagen.challenges(len(steps), 0x5EED0004 + 1)
    d = pc.to_device(pol)
    host = pol.copy()
    # step 0 commits the LDE-domain polynomial before any fold (gen_proof order)
    w0 = 3 << (steps[0] - steps[1])
    tree0 = pc.MerkleTreeGL(3, 2, False, 1 << steps[1], w0)
    root0 = pc.FRI.merkelize(tree0, d, steps[0], steps[1])
    aux0 = oracle.fri_transpose(host, 1 << steps[0], steps[1])
    assert (root0 == oracle.merkletree(aux0.reshape(1 << steps[1], w0), 3)[-4:]).all(), "root 0"
    assert pc.FRI.steps(nbe) == steps
    for s in range(1, len(steps)):
        prev, cur = steps[s - 1], steps[s]
        print("fri step", s, prev, "->", cur, flush=True)
        pc.FRI.fold(s, d, chal[s], nbe, prev, cur)
        host = oracle.fri_fold(host, chal[s], s, nbe, prev, cur)
        _sync()
        got = pc.to_host(d)[:3 << cur]
        assert (got == host).all(), "layer %d" % s
        if s < len(steps) - 1:
            nxt = steps[s + 1]
            w = 3 << (cur - nxt)
            tree = pc.MerkleTreeGL(3, 2, False, 1 << nxt, w)
            root = pc.FRI.merkelize(tree, d, cur, nxt)
            aux = oracle.fri_transpose(host, 1 << cur, nxt)
            expect = oracle.merkletree(aux.reshape(1 << nxt, w), 3)
            assert (root == expect[-4:]).all(), "root %d" % s
            assert (pc.to_host(tree.source) == aux).all()
            # FRI::proveFRIQueries on this layer tree: indices reduced mod 2^current_bits
            queries = [0, 5, (1 << cur) - 1, 123456789, 2**40 + 17]
            proofs = pc.FRI.prove_queries(tree, queries, nxt)
            for qi, q in enumerate(queries):
                exp = oracle.merkle_open(expect, aux.reshape(1 << nxt, w), 3, 2, q % (1 << nxt))
                assert (proofs[qi] == exp).all(), ("fri query", s, q)


def _fri_per_layer(pc, pol_host, chal, nbe, steps):
    """Layer roots and the final polynomial from the per-call entry points."""
    d = pc.to_device(pol_host)
    roots = []
    for s in range(len(steps)):
        if s > 0:
            pc.FRI.fold(s, d, chal[s], nbe, steps[s - 1], steps[s])
        if s < len(steps) - 1:
            t = pc.MerkleTreeGL(3, 2, False, 1 << steps[s + 1], 3 << (steps[s] - steps[s + 1]))
            roots.append(pc.FRI.merkelize(t, d, steps[s], steps[s + 1]))
    _sync()
    return roots, pc.to_host(d)[:3 << steps[-1]]


@pytest.mark.parametrize("nbe", [12, 22])
def test_fri_layers_graph(pc, oracle, nbe):
    """pil2hip_fri_layers (the whole fold + layer-commit loop, graph-captured) equals
    the per-call path: every layer root and the final polynomial, for the graph's
    first run, for replays after new challenges are written into the device
    buffer (eager work and a sync between replays), and for the eager form.
    nbe = 12 is also checked against the oracle."""
    steps = fri_steps(nbe)
    pol = datagen.uniform(3 << nbe, 0x5EED0004 + nbe)
    d_pol = pc.to_device(pol)
    work = pc.empty(3 << steps[1])
    trees = [pc.MerkleTreeGL(3, 2, False, 1 << steps[s + 1], 3 << (steps[s] - steps[s + 1]))
             for s in range(len(steps) - 1)]
    d_ch = pc.empty(3 * len(steps))
    for rep in range(3):
        chal = datagen.challenges(len(steps), 0x5EED0004 + 1 + 17 * rep)
        print("fri layers graph", nbe, "replay", rep, flush=True)
        d_ch.copy_(pc.to_device(np.asarray(chal, dtype=np.uint64).reshape(-1)))
        for t in trees:
            if t.nodes is not None:
                t.nodes.zero_()
        work.zero_()
        _sync()
        pc.FRI.layers(d_pol, work, d_ch, nbe, steps, trees, graph=(rep < 2))
        _sync()
        roots, final = _fri_per_layer(pc, pol, chal, nbe, steps)
        for s, t in enumerate(trees):
            assert (t.get_root() == roots[s]).all(), (rep, "root", s)
        assert (pc.to_host(work)[:3 << steps[-1]] == final).all(), (rep, "final layer")
        if nbe == 12 and rep == 0:
            host = pol.copy()
            for s in range(1, len(steps)):
                host = oracle.fri_fold(host, chal[s], s, nbe, steps[s - 1], steps[s])
            assert (final == host).all()


# ------------------------------------------------------------- openings --
@pytest.mark.parametrize("arity,llv", [(3, 2), (4, 2), (2, 0)])
def test_group_proofs_match_oracle(pc, oracle, arity, llv):
    rows, cols = 1 << 10, 7
    src = datagen.uniform(rows * cols, 99).reshape(rows, cols)
    t = pc.MerkleTreeGL(arity, llv, True, rows, cols, source=pc.to_device(src))
    t.merkelize()
    tree = oracle.merkletree(src, arity)
    for idx in (0, 5, 511, rows - 1):
        assert (t.get_group_proof(idx) == oracle.merkle_open(tree, src, arity, llv, idx)).all()


# ------------------------------------------------------- large (sampled) --
def test_plain_large_commit_sampled(pc, oracle):
    """Plain-large configuration 2^22 x 128, blowup 4, arity 3 (seed 0x5EED0002).
    Size-independent checks: two full columns of the LDE equal the oracle's LDE of
    those columns; sampled leaf digests equal the oracle's sponge over the GPU's
    LDE rows; sampled nodes on every level equal the permutation of their
    children; the root is reproducible."""
    import torch
    nb, nbe, c, arity = 22, 24, 128, 3
    n, ne = 1 << nb, 1 << nbe
    d_trace = pc.empty(n * c)
    pc.fill_uniform(d_trace, 0x5EED0002)
    print("large: commit", flush=True)
    lde, nodes = pc.commit(d_trace, arity, nb, nbe, c)
    root1 = pc.MerkleTreeGL(arity, 0, True, ne, c, nodes=nodes).get_root()
    tr = d_trace.view(n, c)
    for col in (0, 77):
        print("large: column", col, flush=True)
        colv = pc.to_host(tr[:, col].contiguous())
        expect = oracle.lde(colv.reshape(n, 1), ne)[:, 0]
        assert (pc.to_host(lde[col * ne:(col + 1) * ne]) == expect).all(), col
    rng = np.random.default_rng(7)
    rows = np.concatenate([[0, 1, ne - 1], rng.integers(0, ne, 40)])
    L = lde.view(c, ne)
    idx = torch.from_numpy(rows.astype(np.int64)).to(lde.device)
    sampled = pc.to_host(L[:, idx].t().contiguous())
    nd = nodes
    for k, r in enumerate(rows):
        dig = pc.to_host(nd[int(r) * 4:int(r) * 4 + 4])
        assert (dig == oracle.linear_hash(sampled[k], 12)).all(), r
    # every level: sampled parents = permute(children)
    pending, off = ne, 0
    while pending > 1:
        pad = (arity - pending % arity) % arity
        nxt = (pending + arity - 1) // arity
        for i in sorted(set([0, nxt - 1, int(rng.integers(0, nxt))])):
            ch = np.zeros(12, dtype=np.uint64)
            lo = (off + i * arity) * 4
            avail = min(arity, pending - i * arity) * 4
            ch[:avail] = pc.to_host(nd[lo:lo + avail])
            par = pc.to_host(nd[(off + pending + pad + i) * 4:(off + pending + pad + i) * 4 + 4])
            assert (par == oracle.permute(ch)[:4]).all()
        off += pending + pad
        pending = nxt
    lde2, nodes2 = pc.commit(d_trace, arity, nb, nbe, c, lde=lde, nodes=nodes)
    root2 = pc.MerkleTreeGL(arity, 0, True, ne, c, nodes=nodes2).get_root()
    assert (root1 == root2).all()
    # the streamed commit from pinned host memory (eight groups of 16 columns)
    # gives the same LDE and tree, word for word
    print("large: streamed commit from host memory", flush=True)
    h = torch.empty(n * c, dtype=torch.int64, pin_memory=True)
    h.copy_(d_trace.cpu())
    lde3, nodes3 = pc.commit_host(h, arity, nb, nbe, c)
    torch.cuda.synchronize()
    assert torch.equal(nodes3, nodes2)
    assert torch.equal(lde3[:ne * c], lde2[:ne * c])


# -------------------------------------------------------------- grinding --
@pytest.mark.parametrize("ch,nb", [([0x1234567890ABCDEF, 0xFEDCBA0987654321, 0x0123456789ABCDEF], 8), ([1, 2, 3], 1),
                                   ([0, 0, 0], 10), ([P - 1, 2**64 - 1, 12345], 12), ([5, 6, 7], 16)])
def test_grinding_matches_oracle(pc, oracle, ch, nb):
    """Smallest valid nonce (Poseidon2Goldilocks<8>::grinding) equals the oracle's
    exhaustive search; the first two cases are the reference's own test inputs."""
    assert pc.grinding(ch, nb) == oracle.grinding(ch, nb)


def test_grinding_large_matches_compiled_reference(pc):
    """A 20-bit search (about a million permutations, several windows of the
    GPU search) against the reference's own grinding when it is built."""
    from oracle.oracle import RefLib
    if not RefLib.available():
        pytest.skip("oracle/_ref not built")
    ch = [0x0123456789ABCDEF, 42, 2**63 + 5]
    assert pc.grinding(ch, 20) == RefLib().grinding(ch, 20)


# ------------------------------------------------------------ transcript --
@pytest.mark.parametrize("arity", [2, 3, 4])
def test_transcript_matches_oracle(pc, oracle, arity):
    """Device transcript (TranscriptGL) against the oracle on seeded sequences of
    puts, field draws, state reads and query-index draws."""
    from test_oracle import run_transcript, transcript_ops
    for seed in range(3):
        ops = transcript_ops(seed * 10 + arity, 4 * arity)
        t = pc.TranscriptGL(arity)
        assert run_transcript(t, ops) == run_transcript(oracle.transcript(arity), ops), seed
        t.close()


def test_transcript_absorbs_device_root(pc, oracle):
    """put() of a Merkle root that stays in HBM equals put() of its host copy."""
    nb, c = 8, 5
    trace = datagen.uniform((1 << nb) * c, 77).reshape(1 << nb, c)
    lde, nodes = pc.commit(pc.to_device(trace), 3, nb, nb + 1, c)
    root_dev = nodes[-4:]
    t = pc.TranscriptGL(3)
    t.put(root_dev)
    got = t.get_field()
    o = oracle.transcript(3)
    o.put(oracle.merkletree(oracle.lde(trace, 1 << (nb + 1)), 3)[-4:])
    assert (got == o.get_field()).all()
