"""GPU parity tests for the plain transforms (pil2hip_ntt): NTT_Goldilocks::NTT /
INTT (ntt_goldilocks.cpp:211-260) through the C-ABI, against the compiled
reference's fixtures, the oracle and, at 2^24 rows, the round trip. Exact."""
import numpy as np
import pytest

from golden_helpers import colmajor_to_rowmajor, lde_digest
import datagen

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def pc():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    import pil2_commit
    pil2_commit.set_abort_on_error(False)
    return pil2_commit


def _sync():
    import torch
    torch.cuda.synchronize()


def gpu_ntt(pc, a, inverse=False, row_major=True, in_place=False):
    """a: host row-major (size, ncols). Returns the host row-major result."""
    size, ncols = a.shape
    src = pc.to_device(a if row_major else np.ascontiguousarray(a.T))
    dst = None if in_place else pc.empty(size * ncols)
    ntt = pc.NTT_Goldilocks(size)
    (ntt.INTT if inverse else ntt.NTT)(dst, src, size, ncols, row_major=row_major)
    _sync()
    out = pc.to_host(src if in_place else dst)
    if not in_place:
        before = a if row_major else np.ascontiguousarray(a.T)
        assert (pc.to_host(src).reshape(before.shape) == before).all()  # out of place leaves src alone
    return out.reshape(size, ncols) if row_major else colmajor_to_rowmajor(out, size, ncols)


def _oracle_cols(oracle, a, inverse=False):
    f = oracle.intt if inverse else oracle.ntt
    return np.stack([f(a[:, c]) for c in range(a.shape[1])], axis=1)


def test_ntt_matches_compiled_reference(pc, compiled_ref):
    """NTT_Goldilocks::NTT / ::INTT digests of the compiled reference (sizes 1,
    2, 16, 2^10, 2^13, 2^16), row-major as the reference takes them."""
    for case in compiled_ref["ntt"]:
        n, c = 1 << case["log_size"], case["n_cols"]
        print("ntt case", case["log_size"], c, flush=True)
        a = datagen.uniform(n * c, case["seed"]).reshape(n, c)
        assert lde_digest(gpu_ntt(pc, a)) == case["ntt_sha256"], case
        assert lde_digest(gpu_ntt(pc, a, inverse=True)) == case["intt_sha256"], case


@pytest.mark.parametrize("lg,ncols", [(2, 3), (3, 5), (4, 2), (5, 3), (11, 4), (12, 3), (13, 2), (14, 3), (15, 2),
                                      (18, 3), (20, 2)])
@pytest.mark.parametrize("row_major", [True, False])
def test_ntt_matches_oracle(pc, oracle, lg, ncols, row_major):
    """Every plan (direct M < 16, one-block bit reversal M < 2^12, tiled bit
    reversal, with and without the strided pass), both layouts, in place and not."""
    a = datagen.uniform((1 << lg) * ncols, 0xF000 + lg * 16 + ncols).reshape(1 << lg, ncols)
    want_f, want_i = _oracle_cols(oracle, a), _oracle_cols(oracle, a, inverse=True)
    assert (gpu_ntt(pc, a, row_major=row_major) == want_f).all()
    assert (gpu_ntt(pc, a, inverse=True, row_major=row_major) == want_i).all()
    assert (gpu_ntt(pc, a, row_major=row_major, in_place=True) == want_f).all()
    assert (gpu_ntt(pc, a, inverse=True, row_major=row_major, in_place=True) == want_i).all()


def test_ntt_reference_edge_cases(pc):
    """TEST(GOLDILOCKS_TEST, ntt) edge cases, test_ntt_cpu.cpp:139-180: size 1
    is the identity, size 2 round-trips, size 0 and ncols 0 are no-ops."""
    a1 = np.array([[1, 2, 3]], dtype=np.uint64)
    assert (gpu_ntt(pc, a1) == a1).all() and (gpu_ntt(pc, a1, inverse=True) == a1).all()
    a2 = np.array([[1, 2, 3], [4, 5, 6]], dtype=np.uint64)
    assert (gpu_ntt(pc, gpu_ntt(pc, a2), inverse=True) == a2).all()
    x = pc.to_device(a2)
    assert pc.lib.pil2hip_ntt(pc._ptr(x), pc._ptr(x), None, 1, 0, False, True, pc._stream()) == 0
    pc.NTT_Goldilocks(0).NTT(None, x, 0, 3)
    _sync()
    assert (pc.to_host(x).reshape(2, 3) == a2).all()


def test_ntt_full_size_round_trip_and_columns(pc, oracle):
    """2^24 rows x 4 columns (column-major): INTT(NTT(x)) == x everywhere, and
    two whole columns equal the oracle's transform."""
    lg, ncols = 24, 4
    m = 1 << lg
    x = pc.empty(m * ncols)
    pc.fill_uniform(x, 0x5EED0007)
    y = pc.empty(m * ncols)
    z = pc.empty(m * ncols)
    ntt = pc.NTT_Goldilocks(m)
    ntt.NTT(y, x, m, ncols, row_major=False)
    ntt.INTT(z, y, m, ncols, row_major=False)
    _sync()
    import torch
    assert bool(torch.equal(x, z))
    xs = pc.to_host(x[: 2 * m]).reshape(2, m)
    ys = pc.to_host(y[: 2 * m]).reshape(2, m)
    for c in range(2):
        assert (oracle.ntt(xs[c]) == ys[c]).all(), c


def test_ntt_errors(pc):
    x = pc.empty(64 * 2)
    L = pc.lib
    assert L.pil2hip_ntt(pc._ptr(x), pc._ptr(x), None, 25, 1, False, False, pc._stream()) < 0
    assert L.pil2hip_ntt(pc._ptr(x), pc._ptr(x), None, 6, 2, False, True, pc._stream()) < 0  # row-major needs scratch
    assert "scratch" in L.pil2hip_last_error().decode()
