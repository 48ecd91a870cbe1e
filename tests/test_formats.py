"""CPU tests of the on-disk const tree format (host-only entry points of the C-ABI
library; no GPU call): write_const_tree writes the buffer as the file
(MerkleTreeGL::writeFile, merkleTreeGL.cpp:338-345) and load_const_tree reads it
back and checks the root against the verification key JSON
(ConstTree::loadConstTree, const_pols.hpp:66-103)."""
import json

import numpy as np
import pytest

import datagen


@pytest.fixture(scope="module")
def pc():
    import pil2_commit
    pil2_commit.set_abort_on_error(False)
    return pil2_commit


@pytest.mark.parametrize("nbe,nconst,arity", [(3, 2, 2), (10, 5, 3), (12, 1, 4), (5, 0, 3)])
def test_const_tree_size_matches_reference_formula(pc, oracle, nbe, nconst, arity):
    """getConstTreeSizeGL: NExt * nConstants + getNumNodes(NExt)."""
    assert pc.const_tree_size(nbe, nconst, arity) == (1 << nbe) * nconst + oracle.tree_num_elements(1 << nbe, arity)


def test_const_tree_file_round_trip_and_verkey(pc, tmp_path):
    nbe, nconst, arity = 8, 3, 3
    size = pc.const_tree_size(nbe, nconst, arity)
    tree = datagen.uniform(size, 0xC7)
    path = tmp_path / "const.consttree"
    pc.write_const_tree(tree, nbe, nconst, arity, str(path))
    assert path.read_bytes() == tree.astype("<u8").tobytes()
    vk = tmp_path / "verkey.json"
    vk.write_text(json.dumps([int(v) for v in tree[-4:]]))
    ok, back = pc.load_const_tree(str(path), size, str(vk))
    assert ok and (back == tree).all()
    ok, _ = pc.load_const_tree(str(path), size)  # no verification key: load only
    assert ok
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps([int(tree[-4]), int(tree[-3]), int(tree[-2]), int(tree[-1]) ^ 1]))
    assert not pc.load_const_tree(str(path), size, str(bad))[0]
    assert not pc.load_const_tree(str(path), size + 1, str(vk))[0]  # file shorter than the buffer
    assert not pc.load_const_tree(str(tmp_path / "missing"), size, str(vk))[0]
    junk = tmp_path / "junk.json"
    junk.write_text("[1, 2]")
    with pytest.raises(pc.HipError):
        pc.load_const_tree(str(path), size, str(junk))
