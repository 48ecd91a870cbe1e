"""GPU tests of the const tree (ConstTree::calculateConstTreeGL, const_pols.hpp:33-44)
through the C-ABI: the buffer's row-major LDE and nodes against the oracle and the
compiled reference's fixtures, the device entry point against the host one, and
the file round trip with the root checked against a verification key."""
import json

import numpy as np
import pytest

from conftest import hexes
from golden_helpers import colmajor_to_rowmajor, lde_digest
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


@pytest.mark.parametrize("nb,nbe,nconst,arity", [(10, 12, 7, 3), (4, 5, 3, 2), (0, 2, 2, 4), (12, 12, 5, 3)])
def test_const_tree_matches_oracle(pc, oracle, nb, nbe, nconst, arity):
    pols = datagen.uniform((1 << nb) * nconst, 0xA5 + nb).reshape(1 << nb, nconst)
    tree = pc.calculate_const_tree(pols, nb, nbe, arity)
    ne = 1 << nbe
    lde = oracle.lde(pols, ne)
    assert (tree[: ne * nconst].reshape(ne, nconst) == lde).all()
    assert (tree[ne * nconst:] == oracle.merkletree(lde, arity)).all()


def test_const_tree_matches_compiled_reference(pc, compiled_ref):
    """The plain-small fixture shape (2^16 x 32 -> 2^17): LDE digest and roots at
    arity 2/3/4 from the reference's own LDE and merkletree."""
    case = compiled_ref["cases"][0]
    nb, nbe, c = case["n_bits"], case["n_bits_ext"], case["n_cols"]
    pols = datagen.uniform((1 << nb) * c, case["seed"]).reshape(1 << nb, c)
    for a, root in case["roots"].items():
        tree = pc.calculate_const_tree(pols, nb, nbe, int(a))
        assert [int(v) for v in tree[-4:]] == hexes(root), a
    assert lde_digest(tree[: (1 << nbe) * c].reshape(1 << nbe, c)) == case["lde_sha256"]


def test_const_tree_device_entry_and_file(pc, tmp_path):
    import torch
    nb, nbe, nconst, arity = 11, 13, 6, 3
    pols = datagen.uniform((1 << nb) * nconst, 0xC0DE).reshape(1 << nb, nconst)
    host = pc.calculate_const_tree(pols, nb, nbe, arity)
    size = pc.const_tree_size(nbe, nconst, arity)
    d_tree = pc.empty(size)
    d_scratch = pc.empty(((1 << nb) + (1 << nbe)) * nconst)
    rc = pc.lib.pil2hip_const_tree(pc._ptr(d_tree), pc._ptr(pc.to_device(pols)), pc._ptr(d_scratch), nb, nbe,
                                   nconst, arity, pc._stream())
    assert rc == 0
    torch.cuda.synchronize()
    assert (pc.to_host(d_tree) == host).all()
    path, vk = tmp_path / "t.consttree", tmp_path / "t.verkey.json"
    pc.write_const_tree(host, nbe, nconst, arity, str(path))
    vk.write_text(json.dumps([int(v) for v in host[-4:]]))
    ok, back = pc.load_const_tree(str(path), size, str(vk))
    assert ok and (back == host).all()
