"""CPU tests of the C-ABI library: it loads, exports every symbol the public
header declares, keeps the oracle out of the product, and its pure host logic
(tree sizing) agrees with the reference formula. No GPU calls are made."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pil2_commit_hip.h")
LIB = os.path.join(ROOT, "pil2-proofman_amd", "libpil2_commit_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libpil2_commit_hip.so is not built; run __graft_entry__.build()")
    return ctypes.CDLL(LIB)


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref_name in ("set_hash_family", "write_custom_commit", "gen_device_buffers", "free_device_buffers",
                     "get_num_gpus", "set_gpu_mode"):
        assert ref_name in names


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_product_does_not_link_the_oracle():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    assert "oracle_" not in out and "ref_lde" not in out
    deps = subprocess.run(["readelf", "-d", LIB], capture_output=True, text=True).stdout
    assert "liboracle" not in deps and "libref_" not in deps


def test_tree_num_elements_matches_reference_formula(lib, oracle):
    f = lib.pil2hip_tree_num_elements
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    for rows in [1, 2, 3, 4, 5, 8, 9, 27, 100, 1 << 17, (1 << 24) + 3]:
        for arity in (2, 3, 4):
            assert f(rows, arity) == oracle.tree_num_elements(rows, arity)
    # documented sizes: 2^17 rows, arity 3 -> 786,484 words; 2^24 rows -> 100,663,432
    assert f(1 << 17, 3) == 786484
    assert f(1 << 24, 3) == 100663432


def test_set_gpu_mode_reports_gpu_backend(lib):
    f = lib.set_gpu_mode
    f.restype = ctypes.c_bool
    f.argtypes = [ctypes.c_bool]
    assert f(True) is True
    assert f(False) is False


def test_commit_host_stage_words(lib):
    """Staging size of the streamed host commit (pure host logic): three buffers of
    2^nBits rows times the widest group. One group under 256 MiB of trace; else a
    first group of about an eighth of the rate blocks (at least 16 columns) and
    later groups twice as wide (at most the columns left after the first); an
    explicit width keeps every group at that width."""
    f = lib.pil2hip_commit_host_stage_words
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint64] * 4
    assert f(22, 128, 3, 0) == 3 * (1 << 22) * 32      # 4 GiB: 16 columns, then 32
    assert f(16, 32, 3, 0) == 3 * (1 << 16) * 32       # 16 MiB: one group
    assert f(20, 64, 3, 0) == 3 * (1 << 20) * 32       # 512 MiB: 16 columns, then 32
    assert f(22, 300, 4, 0) == 3 * (1 << 22) * 96      # rate 12: 48 columns (4 blocks), then 96
    assert f(22, 128, 2, 0) == 3 * (1 << 22) * 32      # rate 4: 16 columns, then 32
    assert f(22, 24, 3, 0) == 3 * (1 << 22) * 16       # 768 MiB: 16, then the 8 left (under 2 * 16)
    assert f(10, 40, 3, 8) == 3 * (1 << 10) * 8        # explicit width
    assert f(10, 40, 3, 64) == 3 * (1 << 10) * 40      # clamped to the trace
    assert f(10, 0, 3, 0) == 3 * (1 << 10)             # no columns: one word per row
    assert f(10, 40, 5, 0) == 0                        # bad arity


def test_batch_shape_plan_defaults_before_tuning(lib):
    """A shape the batch autotuner has not measured reports 0 and the default plan
    (pure host logic: no device call)."""
    class Plan(ctypes.Structure):
        _fields_ = [("n_streams", ctypes.c_uint32), ("ntt_tile_log", ctypes.c_uint32),
                    ("ntt_radix_log", ctypes.c_uint32), ("coop_max_parents", ctypes.c_uint32)]
    f = lib.pil2hip_batch_shape_plan
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_uint64] * 4 + [ctypes.c_void_p]
    p = Plan()
    assert f(20, 22, 100, 3, ctypes.byref(p)) == 0
    assert (p.n_streams, p.ntt_tile_log, p.ntt_radix_log, p.coop_max_parents) == (0, 13, 4, 16384)
    assert f(20, 22, 100, 3, None) < 0


def test_c_host_links_through_the_header():
    """A plain C program built against include/pil2_commit_hip.h links with the
    library and resolves every entry point it calls (it runs in the GPU tests)."""
    binpath = os.path.join(ROOT, "tests", "c_host", "commit_host")
    if not os.path.exists(binpath):
        pytest.fail("tests/c_host/commit_host is not built; run __graft_entry__.build()")
    out = subprocess.run(["ldd", binpath], capture_output=True, text=True).stdout
    assert "libpil2_commit_hip.so" in out and "not found" not in out, out
    out = subprocess.run([binpath], capture_output=True, text=True)
    assert out.returncode == 2 and "usage" in out.stderr  # no device call without arguments
