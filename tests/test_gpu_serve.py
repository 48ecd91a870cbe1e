"""Commit server (pil2hip_serve_*): jobs admitted by HBM budget while others
run. Every job's root and group proofs must be identical whether it runs alone
or in the mix (exact, integer work), equal the per-instance path, and match
the oracle's Merkle openings on small shapes. Error paths (budget rejection,
bad shapes, cancellation) return codes, never abort."""
import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu
LLV = 2


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


def _mix(n, seed, max_log=13):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        nb = int(rng.integers(3, max_log + 1))
        nbe = nb + int(rng.integers(1, 3))
        nc = int(rng.integers(1, 41))
        q = rng.integers(0, 1 << nbe, 8, dtype=np.uint64)
        jobs.append({"n_bits": nb, "n_bits_ext": nbe, "n_cols": nc, "seed": 9000 + k, "queries": q})
    return jobs


def _drain(eng, n, timeout_polls=600):
    done = {}
    polls = 0
    while len(done) < n:
        for rec in eng.poll(max_n=64, timeout_ms=100):
            done[rec["job_id"]] = rec
        polls += 1
        assert polls < timeout_polls, "server did not complete %d jobs (%d done)" % (n, len(done))
    return done


def _run(pc, jobs, **cfg):
    eng = pc.ServeEngine(**cfg)
    ids = [eng.submit(j["n_bits"], j["n_bits_ext"], j["n_cols"], queries=j["queries"], seed=j["seed"], tag=k)
           for k, j in enumerate(jobs)]
    done = _drain(eng, len(ids))
    out = []
    for jid in ids:
        rec = done[jid]
        assert rec["status"] == 0, (jid, rec["status"])
        out.append((rec["root"].copy(), rec["proofs"][:rec["queries"].size * rec["pw"]].copy()))
    st = eng.stats()
    eng.close()
    return out, st


def test_serve_isolation_alone_vs_mixed(pc):
    """Each job's root and proofs are the same alone (one stream, one slot, one
    job per server) and mixed (many jobs admitted together on 4 streams)."""
    jobs = _mix(24, 1)
    print("serve mix", [(j["n_bits"], j["n_bits_ext"], j["n_cols"]) for j in jobs], flush=True)
    mixed, st = _run(pc, jobs, hbm_budget_bytes=1 << 30, n_streams=4, max_inflight=16, arity=3, llv=LLV)
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["peak_running"] > 1  # the jobs really overlapped
    for k, j in enumerate(jobs):
        alone, _ = _run(pc, [j], hbm_budget_bytes=1 << 28, n_streams=1, max_inflight=1, arity=3, llv=LLV)
        assert np.array_equal(alone[0][0], mixed[k][0]), ("root", k)
        assert np.array_equal(alone[0][1], mixed[k][1]), ("proofs", k)


def test_serve_matches_plain_path_and_oracle(pc, oracle):
    jobs = _mix(10, 2, max_log=10)
    res, _ = _run(pc, jobs, hbm_budget_bytes=1 << 28, n_streams=3, max_inflight=4, arity=3, llv=LLV)
    for k, j in enumerate(jobs):
        nb, nbe, nc = j["n_bits"], j["n_bits_ext"], j["n_cols"]
        trace = datagen.uniform((1 << nb) * nc, j["seed"]).reshape(1 << nb, nc)
        lde = oracle.lde(trace, 1 << nbe)
        tree = oracle.merkletree(lde, 3)
        assert np.array_equal(res[k][0], tree[-4:]), ("root", k)
        pw = pc.proof_words(1 << nbe, nc, 3, LLV)
        for qi, q in enumerate(j["queries"]):
            exp = oracle.merkle_open(tree, lde, 3, LLV, int(q))
            assert exp.size == pw
            assert np.array_equal(res[k][1][qi * pw:(qi + 1) * pw], exp), ("proof", k, qi)
        # the device-resident getGroupProof entry point gives the same proofs
        tr = pc.to_device(trace)
        lde_d, nodes_d = pc.commit(tr, 3, nb, nbe, nc)
        idx = pc.to_device(np.asarray(j["queries"], dtype=np.uint64))
        got = pc.to_host(pc.group_proofs(lde_d, 1 << nbe, nc, nodes_d, 3, LLV, idx))
        _sync()
        assert np.array_equal(got[:len(j["queries"]) * pw], res[k][1]), ("group_proofs", k)


@pytest.mark.parametrize("arity,llv", [(2, 0), (4, 1), (3, 3)])
def test_group_proofs_other_arities(pc, oracle, arity, llv):
    nb, nbe, nc = 7, 9, 11
    trace = datagen.uniform((1 << nb) * nc, 55).reshape(1 << nb, nc)
    lde = oracle.lde(trace, 1 << nbe)
    tree = oracle.merkletree(lde, arity)
    lde_d, nodes_d = pc.commit(pc.to_device(trace), arity, nb, nbe, nc)
    q = np.array([0, 1, 2, 255, 256, 511, 300], dtype=np.uint64)
    got = pc.to_host(pc.group_proofs(lde_d, 1 << nbe, nc, nodes_d, arity, llv, pc.to_device(q)))
    _sync()
    pw = pc.proof_words(1 << nbe, nc, arity, llv)
    for qi, r in enumerate(q):
        assert np.array_equal(got[qi * pw:(qi + 1) * pw], oracle.merkle_open(tree, lde, arity, llv, int(r))), qi


def test_serve_budget_admission(pc):
    """A budget that holds about two jobs at a time: all jobs still complete,
    the arena never exceeds the budget, and results equal a roomy server's."""
    jobs = [{"n_bits": 12, "n_bits_ext": 13, "n_cols": 16, "seed": 500 + k,
             "queries": np.arange(4, dtype=np.uint64) * 1000} for k in range(8)]
    one = pc.serve_job_bytes(12, 13, 16, 3, 4, LLV)
    tight, st = _run(pc, jobs, hbm_budget_bytes=2 * one + 8192, n_streams=4, max_inflight=8, arity=3, llv=LLV)
    assert st["peak_bytes"] <= 2 * one + 8192
    roomy, _ = _run(pc, jobs, hbm_budget_bytes=64 * one, n_streams=4, max_inflight=8, arity=3, llv=LLV)
    for a, b in zip(tight, roomy):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_serve_trace_sources_agree(pc):
    """Host trace, device trace and seeded in-HBM generation of the same values
    give the same root and proofs."""
    nb, nbe, nc, seed = 9, 11, 13, 4242
    trace = datagen.uniform((1 << nb) * nc, seed).reshape(1 << nb, nc)
    dtr = pc.to_device(trace)
    _sync()
    q = [5, 77, 2047]
    eng = pc.ServeEngine(1 << 26, n_streams=2, max_inflight=4, arity=3, llv=LLV)
    ids = [eng.submit(nb, nbe, nc, queries=q, seed=seed), eng.submit(nb, nbe, nc, queries=q, h_trace=trace),
           eng.submit(nb, nbe, nc, queries=q, d_trace=dtr)]
    done = _drain(eng, 3)
    for i in ids[1:]:
        assert done[i]["status"] == 0
        assert np.array_equal(done[i]["root"], done[ids[0]]["root"])
        assert np.array_equal(done[i]["proofs"], done[ids[0]]["proofs"])
    eng.close()


def test_serve_resident_openings_and_release(pc):
    eng = pc.ServeEngine(1 << 26, n_streams=2, max_inflight=4, arity=3, llv=LLV)
    q = np.array([3, 9, 1000, 4095], dtype=np.uint64)
    jid = eng.submit(10, 12, 20, queries=q, seed=31, keep_resident=True)
    rec = _drain(eng, 1)[jid]
    assert rec["status"] == 0
    assert eng.stats()["resident_jobs"] == 1
    again = eng.open(jid, q)
    assert np.array_equal(again, rec["proofs"][:again.size])
    more = eng.open(jid, [4095, 3])
    pw = rec["pw"]
    assert np.array_equal(more[:pw], again[3 * pw:4 * pw]) and np.array_equal(more[pw:], again[:pw])
    eng.release(jid)
    assert eng.stats()["resident_jobs"] == 0 and eng.stats()["bytes_in_use"] == 0
    with pytest.raises(pc.HipError):
        eng.open(jid, q)
    eng.close()


def test_serve_errors_and_cancel(pc):
    eng = pc.ServeEngine(1 << 30, n_streams=1, max_inflight=1, arity=3, llv=LLV)
    with pytest.raises(pc.HipError, match="budget"):
        eng.submit(20, 22, 64)  # 2 GiB of LDE can never fit 1 GiB
    with pytest.raises(pc.HipError):
        eng.submit(10, 9, 4)  # nBitsExt < nBits
    with pytest.raises(pc.HipError):
        eng.submit(4, 25, 4)  # beyond 2^24
    with pytest.raises(pc.HipError):
        eng.submit(4, 5, 4, queries=[32])  # row index out of range
    # one slot: the first (large) job is admitted, the rest wait; cancel the last
    ids = [eng.submit(18, 20, 32, queries=[1], seed=0)] + [eng.submit(12, 14, 8, queries=[1], seed=k)
                                                           for k in range(1, 4)]
    assert eng.cancel(ids[-1]) == 0
    done = _drain(eng, 4)
    assert done[ids[-1]]["status"] == -31
    for i in ids[:-1]:
        assert done[i]["status"] == 0
    assert eng.cancel(ids[0]) < 0  # completed jobs cannot be cancelled
    st = eng.stats()
    assert st["cancelled"] == 1 and st["completed"] == 3
    eng.close()
