"""The library driven from a plain C host (tests/c_host/commit_host.c): the
same C-ABI a Rust FFI shim binds, with no Python in the process. The root it
prints must equal the oracle's root of the same trace."""
import os
import subprocess

import numpy as np
import pytest

import datagen
from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "c_host", "commit_host")


@pytest.mark.parametrize("nb,nbe,cols,arity", [(10, 11, 13, 3), (12, 14, 40, 4), (9, 10, 7, 2)])
def test_c_host_write_custom_commit(oracle, tmp_path, nb, nbe, cols, arity):
    if not os.path.exists(BIN):
        pytest.fail("tests/c_host/commit_host is not built; run __graft_entry__.build()")
    trace = datagen.uniform((1 << nb) * cols, 4000 + nb).reshape(1 << nb, cols)
    path = tmp_path / "trace.bin"
    trace.astype("<u8").tofile(path)
    print("c host", nb, nbe, cols, arity, flush=True)
    out = subprocess.run([BIN, str(path), str(nb), str(nbe), str(cols), str(arity)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    root = [int(w, 16) for w in out.stdout.split()]
    expect = oracle.merkletree(oracle.lde(trace, 1 << nbe), arity)[-4:]
    assert root == [int(v) for v in expect]


@pytest.mark.parametrize("nb,nbe,cols,arity", [(10, 12, 13, 3), (12, 13, 40, 4), (9, 10, 7, 2)])
def test_c_host_commit_witness_matches_compiled_reference(tmp_path, nb, nbe, cols, arity):
    """pil2hip_commit_witness from a plain C host: the LDE and nodes sections it
    leaves in aux_trace (written out by the program) equal the reference's own
    NTT_Goldilocks::LDE and merkletree bytes, and nothing else in aux_trace
    changes (the program checks its sentinels)."""
    from oracle.oracle import Oracle, RefLib
    if not os.path.exists(BIN):
        pytest.fail("tests/c_host/commit_host is not built; run __graft_entry__.build()")
    if not RefLib.available():
        pytest.fail("oracle/_ref (the reference's CPU code) is not built")
    ref = RefLib()
    trace = datagen.uniform((1 << nb) * cols, 4100 + nb).reshape(1 << nb, cols)
    path, out_path = tmp_path / "trace.bin", tmp_path / "aux.bin"
    trace.astype("<u8").tofile(path)
    print("c host witness", nb, nbe, cols, arity, flush=True)
    out = subprocess.run([BIN, "--witness", str(path), str(nb), str(nbe), str(cols), str(arity), str(out_path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = np.fromfile(out_path, dtype="<u8")
    ne = 1 << nbe
    nt = Oracle().tree_num_elements(ne, arity)
    lde = ref.lde(trace, ne)
    tree = ref.merkletree(lde, arity, nt)
    assert got.size == ne * cols + nt
    assert np.array_equal(got[:ne * cols], lde.reshape(-1))
    assert np.array_equal(got[ne * cols:], tree)
    assert [int(w, 16) for w in out.stdout.split()] == [int(v) for v in tree[-4:]]
