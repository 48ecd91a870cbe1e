"""Batched commit (pil2hip_commit_batch): many instances of uneven shapes in one
hipGraph launch. Every root, LDE and tree must equal what the per-instance
path (pil2hip_commit) gives for the same instance alone, and small instances
are also checked against the oracle. Graph replays are checked one by one,
with eager work and synchronisation between them."""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu


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


def _jobs(pc, shapes, seed0):
    jobs = []
    for k, (nb, nbe, nc) in enumerate(shapes):
        tr = pc.empty(max((1 << nb) * nc, 1))
        pc.fill_uniform(tr, seed0 + k)
        jobs.append({"trace": tr, "n_bits": nb, "n_bits_ext": nbe, "n_cols": nc})
    return jobs


def _plain(pc, job, arity):
    lde, nodes = pc.commit(job["trace"], arity, job["n_bits"], job["n_bits_ext"], job["n_cols"])
    _sync()
    return pc.to_host(lde), pc.to_host(nodes)


SHAPES = [(4, 5, 1), (10, 12, 13), (6, 6, 3), (12, 13, 40), (5, 7, 9), (11, 12, 8), (8, 9, 24), (13, 15, 7),
          (9, 10, 1), (7, 9, 17)]


@pytest.mark.parametrize("arity", [2, 3, 4])
def test_batch_equals_plain_and_oracle(pc, oracle, arity):
    print("batch arity", arity, "shapes", SHAPES, flush=True)
    jobs = _jobs(pc, SHAPES, 1000 * arity)
    b = pc.CommitBatch(jobs, arity=arity)
    used = b.launch(n_streams=4)
    assert used == 4
    _sync()
    roots = b.roots_host()
    for k, j in enumerate(b.jobs):
        lde_p, nodes_p = _plain(pc, j, arity)
        assert np.array_equal(pc.to_host(j["lde"]), lde_p), ("lde", k)
        assert np.array_equal(pc.to_host(j["nodes"]), nodes_p), ("nodes", k)
        assert np.array_equal(roots[k], nodes_p[-4:]), ("root", k)
    for k in (0, 2, 4, 8):  # small instances against the oracle
        nb, nbe, nc = SHAPES[k]
        trace = datagen.uniform((1 << nb) * nc, 1000 * arity + k).reshape(1 << nb, nc)
        tree = oracle.merkletree(oracle.lde(trace, 1 << nbe), arity)
        assert np.array_equal(roots[k], tree[-4:]), ("oracle root", k)


def test_batch_graph_replays(pc):
    """The cached graph is replayed; outputs are cleared eagerly between replays
    and each replay must reproduce the first result exactly."""
    jobs = _jobs(pc, SHAPES, 77)
    b = pc.CommitBatch(jobs, arity=3)
    b.launch(n_streams=2)
    _sync()
    ref_roots = b.roots_host().copy()
    ref_nodes = [pc.to_host(j["nodes"]).copy() for j in b.jobs]
    ref_lde = [pc.to_host(j["lde"]).copy() for j in b.jobs]
    for rep in range(4):
        print("replay", rep, flush=True)
        b.roots.zero_()
        for j in b.jobs:
            j["nodes"].zero_()
            j["lde"].fill_(-1)
        _sync()
        assert b.launch(n_streams=2) == 2
        _sync()
        assert np.array_equal(b.roots_host(), ref_roots), rep
        for k, j in enumerate(b.jobs):
            assert np.array_equal(pc.to_host(j["nodes"]), ref_nodes[k]), (rep, k)
            assert np.array_equal(pc.to_host(j["lde"]), ref_lde[k]), (rep, k)


def test_batch_autotune_and_branch_counts(pc):
    jobs = _jobs(pc, SHAPES[:6], 5)
    b = pc.CommitBatch(jobs, arity=3)
    s = b.launch(n_streams=0)
    assert s in (1, 2, 4, 8)  # candidates are capped by the job count
    _sync()
    r0 = b.roots_host().copy()
    assert b.launch(n_streams=0) == s  # cached choice
    _sync()
    assert np.array_equal(b.roots_host(), r0)
    for n in (1, 3, 16):
        b.roots.zero_()
        assert b.launch(n_streams=n) == n
        _sync()
        assert np.array_equal(b.roots_host(), r0), n


def test_batch_mixed_workload_sample(pc):
    """A reduced draw of the mixed-size generator (same rules, smaller bounds)."""
    w = datagen.batched_workload(seed=0x5EED0003, n_instances=24, min_log=10, max_log=15, max_cols=64)
    shapes = [(j["n_bits"], j["n_bits_ext"], j["n_cols"]) for j in w]
    print("mixed sample", shapes, flush=True)
    jobs = _jobs(pc, shapes, 314)
    b = pc.CommitBatch(jobs, arity=3)
    b.launch(n_streams=8)
    _sync()
    roots = b.roots_host()
    for k, j in enumerate(b.jobs):
        _, nodes_p = _plain(pc, j, 3)
        assert np.array_equal(roots[k], nodes_p[-4:]), k


def test_batch_errors(pc):
    jobs = _jobs(pc, SHAPES[:2], 9)
    b = pc.CommitBatch(jobs, arity=3)
    with pytest.raises(pc.HipError):
        b.launch(n_streams=17)
    bad = pc.CommitBatch(jobs, arity=5)
    with pytest.raises(pc.HipError):
        bad.launch(n_streams=1)
    empty = pc.CommitBatch([], arity=3)
    assert empty.launch() == 0


PLAN_SHAPES = SHAPES + [(17, 19, 3), (9, 13, 2), (15, 16, 5)]
PLANS = [dict(n_streams=2, ntt_tile_log=12, ntt_radix_log=2, coop_max_parents=1),
         dict(n_streams=3, ntt_tile_log=13, ntt_radix_log=3, coop_max_parents=4096),
         dict(n_streams=4, ntt_tile_log=14, ntt_radix_log=4, coop_max_parents=65536),
         dict(n_streams=1, ntt_tile_log=12, ntt_radix_log=4, coop_max_parents=16384),
         dict(n_streams=2, ntt_tile_log=13, ntt_radix_log=2, coop_max_parents=1 << 20)]


def test_batch_plans_are_bit_identical(pc, oracle):
    """Every launch plan (NTT tile 2^12..2^14, radix 4/8/16, any cooperative-level
    threshold, any branch count) gives the per-instance path's LDE and tree."""
    jobs = _jobs(pc, PLAN_SHAPES, 4242)
    want = [_plain(pc, j, 3) for j in jobs]
    for plan in PLANS:
        print("plan", plan, flush=True)
        b = pc.CommitBatch(jobs, arity=3)
        assert b.launch(plan=plan) == plan["n_streams"]
        _sync()
        assert b.plan == plan
        roots = b.roots_host()
        for k, j in enumerate(b.jobs):
            assert np.array_equal(pc.to_host(j["lde"]), want[k][0]), (plan, "lde", k)
            assert np.array_equal(pc.to_host(j["nodes"]), want[k][1]), (plan, "nodes", k)
            assert np.array_equal(roots[k], want[k][1][-4:]), (plan, "root", k)
    nb, nbe, nc = PLAN_SHAPES[-3]
    trace = datagen.uniform((1 << nb) * nc, 4242 + len(PLAN_SHAPES) - 3).reshape(1 << nb, nc)
    assert np.array_equal(want[-3][1][-4:], oracle.merkletree(oracle.lde(trace, 1 << nbe), 3)[-4:])


def test_batch_plan_autotune(pc):
    """A plan of all 0 tunes the NTT tile and radix per LDE shape, the level
    threshold per tree shape and the branch count per job list. The outputs equal
    the per-instance path's, every shape gets a valid measured plan, and the tuned
    graph replays from the cache (pil2hip_commit_batch with n_streams = 0 too)."""
    print("autotune shapes", SHAPES, flush=True)
    jobs = _jobs(pc, SHAPES, 99)
    want = [_plain(pc, j, 3) for j in jobs]
    b = pc.CommitBatch(jobs, arity=3)
    s = b.launch(plan=dict(n_streams=0, ntt_tile_log=0, ntt_radix_log=0, coop_max_parents=0))
    _sync()
    assert s >= 1 and b.plan == dict(n_streams=s, ntt_tile_log=0, ntt_radix_log=0, coop_max_parents=0)
    for k, j in enumerate(b.jobs):
        assert np.array_equal(pc.to_host(j["lde"]), want[k][0]), ("lde", k)
        assert np.array_equal(pc.to_host(j["nodes"]), want[k][1]), ("nodes", k)
    for nb, nbe, nc in SHAPES:
        known, p = pc.batch_shape_plan(nb, nbe, nc, 3)
        assert known, (nb, nbe, nc)
        assert 12 <= p["ntt_tile_log"] <= 14 and p["ntt_radix_log"] in (2, 3, 4), p
        assert p["coop_max_parents"] in (4096, 16384, 65536), p
    for rep in range(2):
        for j in b.jobs:
            j["nodes"].zero_()
            j["lde"].zero_()
        _sync()
        assert b.launch(n_streams=0) == s
        _sync()
        for k, j in enumerate(b.jobs):
            assert np.array_equal(pc.to_host(j["lde"]), want[k][0]), (rep, "lde", k)
            assert np.array_equal(pc.to_host(j["nodes"]), want[k][1]), (rep, "nodes", k)
    bad = pc.Pil2CommitPlan(0, 15, 0, 0)
    assert pc.lib.pil2hip_commit_batch_plan(pc.ctypes.cast(b._arr, pc._vp), len(b.jobs), pc.ctypes.byref(bad),
                                            None, pc._ptr(b.roots), pc._stream()) < 0


def test_batch_plan_defaults_and_errors(pc):
    """Fields left 0 (other than the branch count) take the defaults; invalid
    plans are rejected before any capture."""
    jobs = _jobs(pc, SHAPES[:4], 123)
    want = [_plain(pc, j, 3) for j in jobs]
    b = pc.CommitBatch(jobs, arity=3)
    assert b.launch(plan=dict(n_streams=2)) == 2
    _sync()
    assert b.plan == dict(n_streams=2, ntt_tile_log=13, ntt_radix_log=4, coop_max_parents=16384)
    for k, j in enumerate(b.jobs):
        assert np.array_equal(pc.to_host(j["nodes"]), want[k][1]), k
    for bad in (pc.Pil2CommitPlan(2, 15, 0, 0), pc.Pil2CommitPlan(2, 11, 0, 0), pc.Pil2CommitPlan(2, 0, 5, 0),
                pc.Pil2CommitPlan(17, 0, 0, 0)):
        assert pc.lib.pil2hip_commit_batch_plan(pc.ctypes.cast(b._arr, pc._vp), len(b.jobs), pc.ctypes.byref(bad),
                                                None, pc._ptr(b.roots), pc._stream()) < 0
