"""pil2hip_commit_noalloc: the commit whose trace scratch is a device array inside
the library (no allocation by the call). Through the C-ABI, its LDE and tree
must equal pil2hip_commit's (which takes a caller scratch) word for word, on two
calls in a row whose second input is smaller (the shared array keeps the first
call's contents), and on a size at the array's limit; sizes beyond it return
-36 without touching the outputs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORDS = 1 << 24  # PIL2HIP_NOALLOC_SCRATCH_WORDS


@pytest.fixture(scope="module")
def pc():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    import pil2_commit
    pil2_commit.set_hash_family(2)
    pil2_commit.set_abort_on_error(False)
    return pil2_commit


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(pc, nb, nbe, nc, arity, seed):
    import torch
    s = torch.cuda.current_stream()
    st = ctypes.c_void_p(s.cuda_stream)
    n, ne = 1 << nb, 1 << nbe
    trace = pc.empty(max(n * nc, 1))
    pc.fill_uniform(trace, seed)
    tl = pc.get_tree_num_elements(ne, arity)
    lde_a, nodes_a = pc.empty(max(ne * nc, 1)), pc.empty(tl)
    lde_b, nodes_b = pc.empty(max(ne * nc, 1)), pc.empty(tl)
    scratch = pc.empty(max(n * nc, 1))
    lde_b.fill_(-1)
    nodes_b.fill_(-1)
    rc = pc.lib.pil2hip_commit(_vp(lde_a), _vp(nodes_a), _vp(trace), _vp(scratch), arity, nb, nbe, nc, st)
    assert rc == 0, pc.lib.pil2hip_last_error()
    rc = pc.lib.pil2hip_commit_noalloc(_vp(lde_b), _vp(nodes_b), _vp(trace), arity, nb, nbe, nc, st)
    assert rc == 0, pc.lib.pil2hip_last_error()
    torch.cuda.synchronize()
    assert torch.equal(lde_a[:ne * nc], lde_b[:ne * nc]), (nb, nbe, nc)
    assert torch.equal(nodes_a, nodes_b), (nb, nbe, nc)
    return pc.to_host(nodes_b[-4:])


def test_noalloc_equals_commit_twice_second_smaller(pc):
    """A large call then a smaller one (the array still holds the large trace's
    columns), then one at exactly the array's size."""
    print("noalloc: 2^16 x 200 -> 2^18", flush=True)
    r1 = _both(pc, 16, 18, 200, 3, 0x11)
    print("noalloc: 2^9 x 13 -> 2^10", flush=True)
    r2 = _both(pc, 9, 10, 13, 3, 0x12)
    assert not np.array_equal(r1, r2)
    print("noalloc: 2^17 x 128 -> 2^18 (the whole array)", flush=True)
    _both(pc, 17, 18, 128, 4, 0x13)
    print("noalloc: tiny 2^2 x 5 -> 2^4", flush=True)
    _both(pc, 2, 4, 5, 2, 0x14)


def test_noalloc_rejects_oversize(pc):
    import torch
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lde, nodes, trace = pc.empty(8), pc.empty(8), pc.empty(8)
    nodes.fill_(7)
    rc = pc.lib.pil2hip_commit_noalloc(_vp(lde), _vp(nodes), _vp(trace), 3, 17, 18, 129, st)
    assert rc == -36
    assert b"PIL2HIP_NOALLOC_SCRATCH_WORDS" in pc.lib.pil2hip_last_error()
    torch.cuda.synchronize()
    assert (pc.to_host(nodes) == 7).all()
