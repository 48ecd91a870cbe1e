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
