"""The workload configurations at their real sizes, bit for bit against the
reference's own CPU code.

tests/golden/config_reference.json holds what the compiled reference
(NTT_Goldilocks::LDE + Poseidon2Goldilocks<12>::merkletree, arity 3, from
oracle/_ref; tests/golden/make_fixtures.py --configs) gives on the same seeded
inputs:

* plain large: 2^22 x 128 -> 2^24 (seed 0x5EED0002): root, sha256 of the whole
  row-major LDE and of the whole node buffer;
* batched: the 64-instance mixed batch (log2 rows 12..21, 8..256 columns, blowup
  2/4, seed 0x5EED0003): every root, through the autotuned hipGraph batch;
* serving: the first 64 jobs of the Poisson stream (log2 rows 12..22, seed
  0x5EED0005): root and sha256 of the 32 group proofs (llv 2) of each job, served
  in a 200-job mix under the 200 GB budget, and a sample of them served alone.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import datagen
from golden_helpers import device_colmajor_sha256

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_reference.json")


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


def _hexes(lst):
    return [int(x, 16) for x in lst]


def _release():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_plain_large_equals_reference(pc, ref):
    """Root, full row-major LDE digest and full node-buffer digest of the
    2^22 x 128 -> 2^24 commit equal the compiled reference's."""
    import torch
    f = ref["plain_large"]
    nb, nbe, c, arity = f["n_bits"], f["n_bits_ext"], f["n_cols"], f["arity"]
    n, ne = 1 << nb, 1 << nbe
    trace = pc.empty(n * c)
    pc.fill_uniform(trace, f["seed"])
    print("plain_large: commit", flush=True)
    lde, nodes = pc.commit(trace, arity, nb, nbe, c)
    del trace
    root = pc.MerkleTreeGL(arity, 0, True, ne, c, nodes=nodes).get_root()
    assert [int(v) for v in root] == _hexes(f["root"])
    print("plain_large: tree digest", flush=True)
    assert hashlib.sha256(pc.to_host(nodes).astype("<u8").tobytes()).hexdigest() == f["tree_sha256"]
    print("plain_large: LDE digest (row-major, in blocks of 2^20 rows)", flush=True)
    assert device_colmajor_sha256(pc, lde, ne, c) == f["lde_sha256"]
    del lde, nodes
    torch.cuda.synchronize()
    _release()


def test_batched_64_instances_equal_reference(pc, ref):
    """The 64-instance mixed batch through the autotuned hipGraph path: every
    root equals the compiled reference's, on the first (tuning) launch and on a
    replay after the roots were cleared."""
    f = ref["batched"]
    shapes = datagen.batched_workload(seed=f["seed"], n_instances=f["n_instances"])
    assert shapes == f["shapes"]
    jobs = []
    for j in shapes:
        tr = pc.empty((1 << j["n_bits"]) * j["n_cols"])
        pc.fill_uniform(tr, j["seed"])
        jobs.append(dict(j, trace=tr))
    batch = pc.CommitBatch(jobs, arity=f["arity"])
    print("batched: tuning launch", flush=True)
    batch.launch(plan=dict(n_streams=0))
    got = batch.roots_host()
    for k in range(len(shapes)):
        assert [int(v) for v in got[k]] == _hexes(f["roots"][k]), k
    print("batched: replay", flush=True)
    batch.roots.zero_()
    batch.launch(plan=dict(n_streams=0))
    got = batch.roots_host()
    for k in range(len(shapes)):
        assert [int(v) for v in got[k]] == _hexes(f["roots"][k]), k
    del batch, jobs
    _release()


def _serve(pc, eng, shapes, qwords, idxs):
    ids = {}
    for k in idxs:
        j = shapes[k]
        q = qwords[k] % np.uint64(1 << j["n_bits_ext"])
        ids[eng.submit(j["n_bits"], j["n_bits_ext"], j["n_cols"], queries=q, seed=j["seed"], tag=k)] = k
    done, polls = {}, 0
    while len(done) < len(ids):
        for rec in eng.poll(max_n=256, timeout_ms=200):
            assert rec["status"] == 0, (rec["tag"], rec["status"])
            done[ids[rec["job_id"]]] = (rec["root"].copy(), rec["proofs"][:len(rec["queries"]) * rec["pw"]].copy())
            eng.forget(rec["job_id"])
        polls += 1
        assert polls < 3000, "server completed %d of %d jobs" % (len(done), len(ids))
    return done


def _check_job(f, k, root, proofs):
    fj = f["jobs"][k]
    assert [int(v) for v in root] == _hexes(fj["root"]), k
    assert hashlib.sha256(proofs.astype("<u8").tobytes()).hexdigest() == fj["proofs_sha256"], k


def test_serving_prefix_equals_reference(pc, ref):
    """The first 200 jobs of the serving stream, all submitted at once to a server
    with the 200 GB budget (admitted as slots and memory free up): the 64 fixture
    jobs' roots and group proofs equal the reference's. Then a sample of those
    jobs served alone, one at a time, gives the same bytes."""
    import torch
    f = ref["serving"]
    n_jobs = 200
    shapes = datagen.batched_workload(seed=f["seed"], n_instances=n_jobs, min_log=12, max_log=22)
    assert [dict((k, s[k]) for k in ("n_bits", "n_bits_ext", "n_cols", "seed")) for s in shapes[:len(f["jobs"])]] \
        == [dict((k, j[k]) for k in ("n_bits", "n_bits_ext", "n_cols", "seed")) for j in f["jobs"]]
    qwords = datagen.raw_words(n_jobs * f["queries_per_job"], f["query_seed"]).reshape(n_jobs, -1)
    free, _ = torch.cuda.mem_get_info(torch.device("cuda:0"))
    budget = int(min(200e9, free - (8 << 30)))
    print("serving: 200-job mix, budget %.0f GB" % (budget / 1e9), flush=True)
    eng = pc.ServeEngine(budget, n_streams=4, max_inflight=32, arity=f["arity"], llv=f["llv"])
    mixed = _serve(pc, eng, shapes, qwords, range(n_jobs))
    st = eng.stats()
    eng.close()
    assert st["completed"] == n_jobs and st["failed"] == 0
    for k in range(len(f["jobs"])):
        q = qwords[k] % np.uint64(1 << shapes[k]["n_bits_ext"])
        assert [int(v) for v in q] == f["jobs"][k]["queries"], k
        _check_job(f, k, *mixed[k])
    # alone: the smallest and the largest fixture job and two in between
    sizes = [(1 << j["n_bits_ext"]) * j["n_cols"] for j in f["jobs"]]
    order = np.argsort(sizes)
    sample = sorted(set(int(order[i]) for i in (0, len(order) // 3, 2 * len(order) // 3, len(order) - 1)))
    print("serving: alone", sample, flush=True)
    eng = pc.ServeEngine(budget, n_streams=4, max_inflight=32, arity=f["arity"], llv=f["llv"])
    for k in sample:
        alone = _serve(pc, eng, shapes, qwords, [k])
        _check_job(f, k, *alone[k])
        assert np.array_equal(alone[k][1], mixed[k][1])
    eng.close()
    _release()
