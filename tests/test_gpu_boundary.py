"""GPU tests of the drop-in boundary beyond the commit itself, all through the
C-ABI: commit_witness's host-side outputs against the reference's own CPU code,
MerkleTreeGL::getLevel, FRI folds of ratios 2^6..2^10, the batch graph cache
past its eviction point, and the stream contracts of the asynchronous entry
points. Integer work: every comparison is exact."""
import ctypes

import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001
SENTINEL = np.uint64(0x5A5A5A5A5A5A5A5A)


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
    from oracle.oracle import RefLib
    if not RefLib.available():
        pytest.fail("oracle/_ref (the reference's CPU code) is not built")
    return RefLib()


def _sync():
    import torch
    torch.cuda.synchronize()


# ------------------------------------------------------------ commit_witness --
@pytest.mark.parametrize("nb,nbe,cols,arity", [(10, 12, 13, 3), (12, 13, 40, 4), (9, 10, 7, 2), (14, 16, 64, 3),
                                               (4, 5, 3, 3), (0, 1, 2, 3)])
def test_commit_witness_matches_compiled_reference(pc, ref, oracle, nb, nbe, cols, arity):
    """commit_witness_cpu (starks_api.cpp:683-723) leaves the row-major LDE at
    aux_trace[offset_dst], the nodes at aux_trace[offset_mt] and the root; the
    reference's own NTT_Goldilocks::LDE and merkletree on the same trace give the
    expected bytes. Every other word of aux_trace keeps its sentinel."""
    n, ne = 1 << nb, 1 << nbe
    trace = datagen.uniform(n * cols, 7000 + nb * 8 + arity).reshape(n, cols)
    nt = oracle.tree_num_elements(ne, arity)
    off_dst, off_mt = n * cols + 11, n * cols + 11 + ne * cols + 5
    aux = np.full(off_mt + nt + 9, SENTINEL, dtype=np.uint64)
    aux[:n * cols] = trace.reshape(-1)
    root = np.zeros(4, dtype=np.uint64)
    print("commit_witness", nb, nbe, cols, arity, flush=True)
    pc.commit_witness(root, arity, nb, nbe, cols, trace, aux, off_dst, off_mt)
    lde = ref.lde(trace, ne)
    tree = ref.merkletree(lde, arity, nt)
    assert np.array_equal(aux[off_dst:off_dst + ne * cols], lde.reshape(-1))
    assert np.array_equal(aux[off_mt:off_mt + nt], tree)
    assert np.array_equal(root, tree[-4:])
    assert np.array_equal(aux[:n * cols], trace.reshape(-1))
    assert (aux[n * cols:off_dst] == SENTINEL).all() and (aux[off_dst + ne * cols:off_mt] == SENTINEL).all()
    assert (aux[off_mt + nt:] == SENTINEL).all()


def test_commit_witness_with_device_buffers_and_errors(pc, ref, oracle):
    nb, nbe, cols, arity = 11, 12, 9, 3
    n, ne = 1 << nb, 1 << nbe
    trace = datagen.uniform(n * cols, 7100).reshape(n, cols)
    nt = oracle.tree_num_elements(ne, arity)
    h = pc.gen_device_buffers(arity, nbe)
    aux = np.zeros(ne * cols + nt, dtype=np.uint64)
    root = np.zeros(4, dtype=np.uint64)
    pc.commit_witness(root, arity, nb, nbe, cols, trace, aux, 0, ne * cols, d_buffers=h)
    pc.free_device_buffers(h)
    assert np.array_equal(root, ref.merkletree(ref.lde(trace, ne), arity, nt)[-4:])
    ptr = aux.ctypes.data_as(ctypes.c_void_p)
    # overlapping sections, bad arity, nBitsExt < nBits: error codes, no abort
    rc = pc.lib.pil2hip_commit_witness(root.ctypes.data_as(ctypes.c_void_p), arity, nb, nbe, cols,
                                       trace.ctypes.data_as(ctypes.c_void_p), ptr, 0, 8, None)
    assert rc == -35 and b"overlap" in pc.lib.pil2hip_last_error()
    assert pc.lib.pil2hip_commit_witness(root.ctypes.data_as(ctypes.c_void_p), 5, nb, nbe, cols,
                                         trace.ctypes.data_as(ctypes.c_void_p), ptr, 0, ne * cols, None) < 0
    assert pc.lib.pil2hip_commit_witness(root.ctypes.data_as(ctypes.c_void_p), arity, nb, nb - 1, cols,
                                         trace.ctypes.data_as(ctypes.c_void_p), ptr, 0, ne * cols, None) < 0


# ------------------------------------------------------------------ getLevel --
def _level_reference(tree, height, arity, llv):
    """merkleTreeGL.cpp:84-101 restated: the first level with <= arity^llv nodes."""
    lim = arity ** llv
    n, off = height, 0
    while n > lim:
        n = (n + arity - 1) // arity
        off += n * arity * 4
    out = np.zeros(lim * 4, dtype=np.uint64)
    out[:n * 4] = tree[off:off + n * 4]
    return out, n


def _reduce(oracle, digests, n, arity):
    """Hash n digests up to one root: levels zero-padded to a multiple of arity,
    parent = first 4 words of permute(children) (poseidon2_goldilocks.cpp:130-175)."""
    cur = [digests[4 * i:4 * i + 4] for i in range(n)]
    while len(cur) > 1:
        cur = cur + [np.zeros(4, dtype=np.uint64)] * ((arity - len(cur) % arity) % arity)
        cur = [oracle.permute(np.concatenate(cur[i:i + arity]))[:4] for i in range(0, len(cur), arity)]
    return cur[0]


@pytest.mark.parametrize("arity", [2, 3, 4])
@pytest.mark.parametrize("rows", [1, 5, 64, 1000, 4096])
def test_get_level_matches_tree_and_rehashes_to_root(pc, oracle, arity, rows):
    """getLevel at llv 0..3: the words equal the tree level the reference copies,
    zero-padded to arity^llv digests, and hashing that level up (the verifier's
    use of last_levels) gives the root, which pins the level's offset."""
    cols = 5
    src = datagen.uniform(rows * cols, 8000 + rows).reshape(rows, cols)
    tree = oracle.merkletree(src, arity)
    t = pc.MerkleTreeGL(arity, 0, False, rows, cols, source=pc.to_device(src))
    t.merkelize()
    _sync()
    assert np.array_equal(pc.to_host(t.nodes), tree)
    for llv in range(4):
        t.last_level_verification = llv
        got = t.get_level()
        if llv == 0:
            assert got.size == 0
            continue
        want, n = _level_reference(tree, rows, arity, llv)
        assert got.size == 4 * arity ** llv and np.array_equal(got, want), llv
        assert np.array_equal(_reduce(oracle, got, n, arity), tree[-4:]), llv
    assert pc.lib.pil2hip_level_words(arity, 0) == 0 and pc.lib.pil2hip_level_words(arity, 2) == 4 * arity ** 2


# -------------------------------------------------------- large FRI folds --
@pytest.mark.parametrize("prev,cur,nbe", [(12, 6, 14), (13, 6, 13), (14, 6, 16), (15, 6, 15), (16, 6, 16),
                                          (10, 0, 12), (11, 3, 11)])
def test_fri_fold_large_ratios_match_oracle(pc, oracle, prev, cur, nbe):
    """FRI::fold (fri.hpp:32-99) at ratios 2^6..2^10 (two folds in one kernel),
    in place and out of place, against the oracle's direct fold of that ratio."""
    pol = datagen.uniform(3 << prev, 0x5EED0100 + prev * 16 + cur)
    ch = datagen.uniform(3, 0x5EED0200 + prev)
    want = oracle.fri_fold(pol, ch, 1, nbe, prev, cur)
    print("fri fold", prev, cur, nbe, flush=True)
    d = pc.to_device(pol)
    out = pc.empty(3 << cur)
    pc.FRI.fold(1, d, ch, nbe, prev, cur, out=out)
    _sync()
    assert np.array_equal(pc.to_host(out), want)
    pc.FRI.fold(1, d, ch, nbe, prev, cur)  # in place
    _sync()
    assert np.array_equal(pc.to_host(d)[:3 << cur], want)


@pytest.mark.parametrize("prev,cur", [(25, 22), (25, 19)])
def test_fri_fold_above_2_24_matches_oracle(pc, oracle, prev, cur):
    """Folds of a 2^25-point layer: the lanes' coset roots w_{2^25}^-g come from the
    2^28 domain's split root tables (the 2^24 ones cover layers up to 2^24 points);
    ratio 64 also takes the second fold's roots from the 2^24 tables."""
    pol = datagen.uniform(3 << prev, 0x5EED0300 + prev * 16 + cur)
    ch = datagen.uniform(3, 0x5EED0400 + cur)
    want = oracle.fri_fold(pol, ch, 1, prev, prev, cur)
    print("fri fold", prev, cur, flush=True)
    d = pc.to_device(pol)
    out = pc.empty(3 << cur)
    pc.FRI.fold(1, d, ch, prev, prev, cur, out=out)
    _sync()
    assert np.array_equal(pc.to_host(out), want)


def test_fri_fold_ratio_limit(pc):
    d = pc.empty(3 << 12)
    ch = np.array([1, 2, 3], dtype=np.uint64)
    rc = pc.lib.pil2hip_fri_fold(pc._ptr(d), pc._ptr(d), ch.ctypes.data_as(ctypes.c_void_p), 1, 12, 12, 1,
                                 pc._stream())
    assert rc == -17


def test_fri_layers_with_large_steps(pc, oracle):
    """The FRI layer loop with steps that drop 7 and 6 bits (graph and eager)
    against the oracle's folds and trees."""
    nbe, steps = 14, [14, 7, 1]
    pol = datagen.uniform(3 << nbe, 0x5EED0300)
    chal = datagen.challenges(len(steps), 0x5EED0301)
    host, roots = pol.copy(), []
    for s in range(len(steps)):
        if s > 0:
            host = oracle.fri_fold(host, chal[s], s, nbe, steps[s - 1], steps[s])
        if s + 1 < len(steps):
            src = oracle.fri_transpose(host, 1 << steps[s], steps[s + 1])
            rows, cols = 1 << steps[s + 1], 3 << (steps[s] - steps[s + 1])
            roots.append(oracle.merkletree(src.reshape(rows, cols), 3)[-4:])
    d_pol = pc.to_device(pol)
    d_ch = pc.to_device(np.asarray(chal, dtype=np.uint64).reshape(-1))
    for graph in (True, False):
        work = pc.empty(3 << steps[1])
        trees = [pc.MerkleTreeGL(3, 2, False, 1 << steps[s + 1], 3 << (steps[s] - steps[s + 1]))
                 for s in range(len(steps) - 1)]
        pc.FRI.layers(d_pol, work, d_ch, nbe, steps, trees, graph=graph)
        _sync()
        for s, t in enumerate(trees):
            assert np.array_equal(t.get_root(), roots[s]), (graph, s)
        assert np.array_equal(pc.to_host(work)[:3 << steps[-1]], host), graph


# ------------------------------------------------------ batch graph cache --
def test_batch_graph_cache_past_eviction(pc):
    """More than 8 distinct job lists launched back to back without a sync: the
    cache evicts its oldest graphs while later ones are still queued, and every
    batch's roots equal the per-instance path."""
    shapes = [(8, 9, 5), (10, 11, 12), (9, 11, 3)]
    batches, want = [], []
    for b in range(11):
        jobs = []
        for k, (nb, nbe, nc) in enumerate(shapes):
            tr = pc.empty((1 << nb) * nc)
            pc.fill_uniform(tr, 9100 + 10 * b + k)
            jobs.append({"trace": tr, "n_bits": nb, "n_bits_ext": nbe, "n_cols": nc})
        batches.append(pc.CommitBatch(jobs, arity=3))
    _sync()
    for b in batches:
        want.append([pc.to_host(pc.commit(j["trace"], 3, j["n_bits"], j["n_bits_ext"], j["n_cols"])[1])[-4:]
                     for j in b.jobs])
    _sync()
    print("launching 11 batches twice", flush=True)
    for rep in range(2):
        for b in batches:
            b.roots.zero_()
            assert b.launch(n_streams=2) == 2
        _sync()
        for i, b in enumerate(batches):
            roots = b.roots_host()
            for k in range(len(shapes)):
                assert np.array_equal(roots[k], want[i][k]), (rep, i, k)
    pc.batch_release()


# ---------------------------------------------------------- stream contracts --
def test_commit_host_two_streams_share_stage(pc):
    """Two streamed host commits on two streams share one staging buffer. The
    stage is in use until the first commit's stream has run its last LDE, so the
    second stream waits for an event recorded after the first commit; both LDEs
    and trees equal pil2hip_commit's."""
    import torch
    nb, nbe, cols, arity, gc = 14, 15, 64, 3, 16
    n = 1 << nb
    traces = [datagen.uniform(n * cols, 9200 + k).reshape(n, cols) for k in range(2)]
    pinned = [torch.from_numpy(t.view(np.int64).reshape(-1).copy()).pin_memory() for t in traces]
    words = pc.lib.pil2hip_commit_host_stage_words(nb, cols, arity, gc)
    stage = pc.empty(words)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    _sync()
    out1 = pc.commit_host(pinned[0], arity, nb, nbe, cols, stage=stage, group_cols=gc, stream=s1)
    ev = torch.cuda.Event()
    ev.record(s1)
    s2.wait_event(ev)
    out2 = pc.commit_host(pinned[1], arity, nb, nbe, cols, stage=stage, group_cols=gc, stream=s2)
    _sync()
    for k, (lde, nodes) in enumerate((out1, out2)):
        lde_d, nodes_d = pc.commit(pc.to_device(traces[k]), arity, nb, nbe, cols)
        _sync()
        assert np.array_equal(pc.to_host(lde), pc.to_host(lde_d)), k
        assert np.array_equal(pc.to_host(nodes), pc.to_host(nodes_d)), k


def test_wrapper_temporaries_on_a_side_stream(pc):
    """The Python wrappers allocate their temporaries (scratch, staging, outputs)
    in the context of the stream they launch on, so the caching allocator cannot
    hand those blocks to other work while the kernels still use them. Commits on
    a side stream are followed at once by allocations and writes on the current
    stream; every root must still be right."""
    import torch
    nb, nbe, cols, arity = 13, 15, 48, 3
    traces = [datagen.uniform((1 << nb) * cols, 9300 + k).reshape(1 << nb, cols) for k in range(4)]
    want = []
    for t in traces:
        _, nodes = pc.commit(pc.to_device(t), arity, nb, nbe, cols)
        _sync()
        want.append(pc.to_host(nodes)[-4:].copy())
    side = torch.cuda.Stream()
    d_tr = [pc.to_device(t) for t in traces]
    _sync()
    roots = []
    for k in range(4):
        _, nodes = pc.commit(d_tr[k], arity, nb, nbe, cols, stream=side)
        root = pc.empty(4, stream=side)
        with torch.cuda.stream(side):
            root.copy_(nodes[-4:])
        roots.append(root)
        del nodes
        junk = [pc.empty((1 << nbe) * cols) for _ in range(2)]  # on the current stream
        for j in junk:
            j.fill_(-1)
    _sync()
    for k in range(4):
        assert np.array_equal(pc.to_host(roots[k]), want[k]), k
