"""Concurrent calls into the C-ABI from several host threads.

proofman commits instances from concurrent contribution threads
(commit_witness, starks_api.hpp:98, called from proofman.rs:2495-2520), so the
entry points must be re-entrant per call. Each thread here commits its own
instances on its own HIP stream while the others do the same; every LDE, tree
and root must equal the result of the same commit run alone. ctypes releases
the GIL around foreign calls, so the library really is entered concurrently."""
import threading

import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

SHAPES = [(10, 12, 13), (12, 13, 40), (8, 9, 24), (11, 12, 8), (13, 15, 7), (9, 11, 33)]


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


def _run_threads(target, n):
    errors = []

    def wrap(k):
        try:
            target(k)
        except BaseException as e:  # reported by the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=wrap, args=(k,)) for k in range(n)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a commit thread did not finish"
    assert not errors, errors


def test_concurrent_device_commits(pc):
    import torch
    print("threads: device commits", SHAPES, flush=True)
    jobs = []
    for k, (nb, nbe, nc) in enumerate(SHAPES):
        tr = pc.empty((1 << nb) * nc)
        pc.fill_uniform(tr, 777 + k)
        jobs.append(dict(trace=tr, nb=nb, nbe=nbe, nc=nc))
    _sync()
    want = []
    for j in jobs:
        lde, nodes = pc.commit(j["trace"], 3, j["nb"], j["nbe"], j["nc"])
        _sync()
        want.append((pc.to_host(lde), pc.to_host(nodes)))
    for j in jobs:
        j["lde"] = pc.empty((1 << j["nbe"]) * j["nc"])
        j["nodes"] = pc.empty(pc.get_tree_num_elements(1 << j["nbe"], 3))
        j["scratch"] = pc.empty((1 << j["nb"]) * j["nc"])
        j["stream"] = torch.cuda.Stream()
    _sync()
    got = [[] for _ in jobs]

    def work(k):
        j = jobs[k]
        for rep in range(3):
            j["nodes"].zero_()  # on the current (default) stream, ordered before the commit below
            _sync()
            pc.commit(j["trace"], 3, j["nb"], j["nbe"], j["nc"], lde=j["lde"], nodes=j["nodes"],
                      scratch=j["scratch"], stream=j["stream"])
            j["stream"].synchronize()
            got[k].append((pc.to_host(j["lde"]), pc.to_host(j["nodes"])))

    _run_threads(work, len(jobs))
    for k in range(len(jobs)):
        for rep, (lde, nodes) in enumerate(got[k]):
            assert np.array_equal(lde, want[k][0]), (k, rep, "lde")
            assert np.array_equal(nodes, want[k][1]), (k, rep, "nodes")


def test_concurrent_write_custom_commit(pc):
    """write_custom_commit (host trace in, library-internal device buffers) from
    several threads at once gives each thread the root of its own trace."""
    print("threads: write_custom_commit", SHAPES, flush=True)
    traces = [datagen.uniform((1 << nb) * nc, 900 + k).reshape(1 << nb, nc) for k, (nb, nbe, nc) in enumerate(SHAPES)]
    want = []
    for (nb, nbe, nc), tr in zip(SHAPES, traces):
        root = np.zeros(4, dtype=np.uint64)
        pc.write_custom_commit(root, 3, nb, nbe, nc, None, tr, "")
        want.append(root.copy())
    got = [[] for _ in SHAPES]

    def work(k):
        nb, nbe, nc = SHAPES[k]
        for rep in range(3):
            root = np.zeros(4, dtype=np.uint64)
            pc.write_custom_commit(root, 3, nb, nbe, nc, None, traces[k], "")
            got[k].append(root)

    _run_threads(work, len(SHAPES))
    for k in range(len(SHAPES)):
        for rep, r in enumerate(got[k]):
            assert np.array_equal(r, want[k]), (k, rep)
