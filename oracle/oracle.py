"""ctypes loader for the oracle libraries (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker or the timed CPU baseline. The product library
(libpil2_commit_hip.so) never loads anything from oracle/.

Two libraries are exposed:
  * ``Oracle`` — our C restatement (gl_oracle.c), always available once built.
  * ``RefLib`` — the reference's own CPU sources compiled by oracle/Makefile
    (ntt_goldilocks.cpp, poseidon2_goldilocks.cpp, ...); present where it was
    built. The AVX-512 build is used only when the host CPU reports avx512f.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
P = 0xFFFFFFFF00000001

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


def build(quiet=True):
    """Compile the restatement (and the reference libraries where its sources exist)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def _host_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


class Oracle:
    """Our C restatement of the commitment path."""

    def __init__(self, path=None):
        path = path or os.path.join(REF_DIR, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        self.L = L
        for name in ("oracle_gl_add", "oracle_gl_sub", "oracle_gl_mul", "oracle_gl_pow"):
            getattr(L, name).restype = _u64
            getattr(L, name).argtypes = [_u64, _u64]
        L.oracle_gl_inv.restype = _u64
        L.oracle_gl_inv.argtypes = [_u64]
        L.oracle_gl_root.restype = _u64
        L.oracle_gl_root.argtypes = [_u32]
        L.oracle_c3_mul.argtypes = [_u64p, _u64p, _u64p]
        L.oracle_ntt.argtypes = [_u64p, _u64]
        L.oracle_intt.argtypes = [_u64p, _u64]
        L.oracle_dft_naive.argtypes = [_u64p, _u64p, _u64, ctypes.c_int]
        L.oracle_poly_eval.restype = _u64
        L.oracle_poly_eval.argtypes = [_u64p, _u64, _u64]
        L.oracle_lde.argtypes = [_u64p, _u64p, _u64, _u64, _u64]
        L.oracle_compute_q.argtypes = [_u64p, _u64p, _u64, _u64, _u64, _u64]
        L.oracle_p2_permute.argtypes = [_u64p, _u32]
        L.oracle_p2_linear_hash.argtypes = [_u64p, _u64p, _u64, _u64, _u32]
        L.oracle_tree_num_elements.restype = _u64
        L.oracle_tree_num_elements.argtypes = [_u64, _u64]
        L.oracle_tr_size.restype = _u64
        L.oracle_tr_init.argtypes = [ctypes.c_void_p, _u32]
        L.oracle_tr_put.argtypes = [ctypes.c_void_p, _u64p, _u64]
        L.oracle_tr_get_field.argtypes = [ctypes.c_void_p, _u64p]
        L.oracle_tr_get_state.argtypes = [ctypes.c_void_p, _u64p, _u32]
        L.oracle_tr_get_permutations.argtypes = [ctypes.c_void_p, _u64p, _u64, _u64]
        L.oracle_p2_grinding.restype = _u64
        L.oracle_p2_grinding.argtypes = [_u64p, _u32, _u64]
        L.oracle_p2_merkletree.argtypes = [_u64p, _u64p, _u64, _u64, _u64, _u64, _u64]
        L.oracle_merkle_open.restype = _u64
        L.oracle_merkle_open.argtypes = [_u64p, _u64p, _u64p, _u64, _u64, _u64, _u64, _u64, _u64, _u64]
        L.oracle_fri_fold.restype = ctypes.c_int
        L.oracle_fri_fold.argtypes = [_u64p, _u64p, _u64p, _u64, _u64, _u64, _u64]
        L.oracle_fri_transpose.argtypes = [_u64p, _u64p, _u64, _u64]
        L.oracle_fri_fold_one.argtypes = [_u64p, _u64p, _u64p, _u64, _u64, _u64, _u64]
        L.oracle_fill_uniform.argtypes = [_u64p, _u64, _u64]
        L.oracle_num_threads.restype = ctypes.c_int

    # field -------------------------------------------------------------
    def add(self, a, b): return self.L.oracle_gl_add(a, b)
    def sub(self, a, b): return self.L.oracle_gl_sub(a, b)
    def mul(self, a, b): return self.L.oracle_gl_mul(a, b)
    def pow(self, a, e): return self.L.oracle_gl_pow(a, e)
    def inv(self, a): return self.L.oracle_gl_inv(a)
    def root(self, k): return self.L.oracle_gl_root(k)

    def c3_mul(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        r = np.zeros(3, dtype=np.uint64)
        self.L.oracle_c3_mul(_ptr(r), _ptr(a), _ptr(b))
        return r

    # NTT ---------------------------------------------------------------
    def ntt(self, a):
        a = np.array(a, dtype=np.uint64)
        self.L.oracle_ntt(_ptr(a), a.size)
        return a

    def intt(self, a):
        a = np.array(a, dtype=np.uint64)
        self.L.oracle_intt(_ptr(a), a.size)
        return a

    def dft_naive(self, a, inverse=False):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = np.zeros_like(a)
        self.L.oracle_dft_naive(_ptr(out), _ptr(a), a.size, int(inverse))
        return out

    def poly_eval(self, coeffs, x):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        return self.L.oracle_poly_eval(_ptr(c), c.size, x)

    def lde(self, trace, n_ext):
        """trace: row-major (N, ncols) uint64 -> row-major (n_ext, ncols)."""
        trace = np.ascontiguousarray(trace, dtype=np.uint64)
        n, ncols = trace.shape
        out = np.zeros((n_ext, ncols), dtype=np.uint64)
        self.L.oracle_lde(_ptr(out), _ptr(trace), n_ext, n, ncols)
        return out

    def compute_q(self, q, n, q_deg):
        """q: row-major (NExt, qDim) -> cmQ row-major (NExt, q_deg*qDim)."""
        q = np.ascontiguousarray(q, dtype=np.uint64)
        n_ext, q_dim = q.shape
        out = np.zeros((n_ext, q_deg * q_dim), dtype=np.uint64)
        self.L.oracle_compute_q(_ptr(out), _ptr(q), n, n_ext, q_deg, q_dim)
        return out

    # Poseidon2 ---------------------------------------------------------
    def permute(self, state):
        st = np.array(state, dtype=np.uint64)
        self.L.oracle_p2_permute(_ptr(st), st.size)
        return st

    def linear_hash(self, vals, width):
        v = np.ascontiguousarray(vals, dtype=np.uint64)
        out = np.zeros(4, dtype=np.uint64)
        self.L.oracle_p2_linear_hash(_ptr(out), _ptr(v) if v.size else _ptr(np.zeros(1, np.uint64)), v.size, 1, width)
        return out

    def transcript(self, arity):
        return OracleTranscript(self.L, arity)

    def grinding(self, challenge, n_bits, limit=1 << 24):
        ch = np.ascontiguousarray(np.asarray(challenge, dtype=np.uint64)[:3])
        return int(self.L.oracle_p2_grinding(_ptr(ch), n_bits, limit))

    def tree_num_elements(self, rows, arity):
        return self.L.oracle_tree_num_elements(rows, arity)

    def merkletree(self, src, arity):
        """src: row-major (rows, cols) -> full tree buffer (reference layout)."""
        src = np.ascontiguousarray(src, dtype=np.uint64)
        rows, cols = src.shape
        tree = np.zeros(self.tree_num_elements(rows, arity), dtype=np.uint64)
        self.L.oracle_p2_merkletree(_ptr(tree), _ptr(src), cols, rows, cols, 1, arity)
        return tree

    def merkle_open(self, tree, src, arity, llv, idx):
        src = np.ascontiguousarray(src, dtype=np.uint64)
        rows, cols = src.shape
        buf = np.zeros(cols + 64 * 4 * arity, dtype=np.uint64)
        n = self.L.oracle_merkle_open(_ptr(buf), _ptr(tree), _ptr(src), cols, rows, cols, 1, arity, llv, idx)
        return buf[:n]

    # FRI ---------------------------------------------------------------
    def fri_fold(self, pol, challenge, step, n_bits_ext, prev_bits, current_bits):
        pol = np.ascontiguousarray(pol, dtype=np.uint64)
        ch = np.ascontiguousarray(challenge, dtype=np.uint64)
        out = np.zeros(3 << current_bits, dtype=np.uint64)
        self.L.oracle_fri_fold(_ptr(out), _ptr(pol), _ptr(ch), step, n_bits_ext, prev_bits, current_bits)
        return out

    def fri_fold_one(self, vals, challenge, n_bits_ext, prev_bits, current_bits, g):
        v = np.ascontiguousarray(vals, dtype=np.uint64)
        ch = np.ascontiguousarray(challenge, dtype=np.uint64)
        out = np.zeros(3, dtype=np.uint64)
        self.L.oracle_fri_fold_one(_ptr(out), _ptr(v), _ptr(ch), n_bits_ext, prev_bits, current_bits, g)
        return out

    def fri_transpose(self, pol, degree, t_bits):
        pol = np.ascontiguousarray(pol, dtype=np.uint64)
        aux = np.zeros_like(pol)
        self.L.oracle_fri_transpose(_ptr(aux), _ptr(pol), degree, t_bits)
        return aux

    # data --------------------------------------------------------------
    def uniform(self, n, seed):
        out = np.zeros(n, dtype=np.uint64)
        self.L.oracle_fill_uniform(_ptr(out), n, seed)
        return out

    def num_threads(self):
        return self.L.oracle_num_threads()


class OracleTranscript:
    """TranscriptGL restated in C (oracle_tr_*)."""

    def __init__(self, L, arity):
        self.L = L
        self.buf = ctypes.create_string_buffer(int(L.oracle_tr_size()))
        L.oracle_tr_init(self.buf, arity)

    def put(self, vals):
        v = np.ascontiguousarray(np.asarray(vals, dtype=np.uint64))
        if v.size:
            self.L.oracle_tr_put(self.buf, _ptr(v), v.size)

    def get_field(self):
        out = np.zeros(3, dtype=np.uint64)
        self.L.oracle_tr_get_field(self.buf, _ptr(out))
        return out

    def get_state(self, n=4):
        out = np.zeros(n, dtype=np.uint64)
        self.L.oracle_tr_get_state(self.buf, _ptr(out), n)
        return out

    def get_permutations(self, n, n_bits):
        out = np.zeros(n, dtype=np.uint64)
        self.L.oracle_tr_get_permutations(self.buf, _ptr(out), n, n_bits)
        return out


class RefLib:
    """The reference's own CPU code, compiled by oracle/Makefile."""

    def __init__(self, path=None):
        if path is None:
            name = "libref_avx512.so" if _host_has_avx512() else "libref_avx2.so"
            path = os.path.join(REF_DIR, name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        L = ctypes.CDLL(path)
        self.L = L
        L.ref_lde.argtypes = [_u64p, _u64p, _u64, _u64, _u64]
        L.ref_ntt.argtypes = [_u64p, _u64, _u64, ctypes.c_int]
        L.ref_compute_q.argtypes = [_u64p, _u64p, _u64, _u64, _u64, _u64]
        L.ref_merkletree.restype = ctypes.c_int
        L.ref_merkletree.argtypes = [_u64p, _u64p, _u64, _u64, _u64]
        L.ref_permute.restype = ctypes.c_int
        L.ref_permute.argtypes = [_u64p, _u32]
        L.ref_linear_hash.restype = ctypes.c_int
        L.ref_linear_hash.argtypes = [_u64p, _u64p, _u64, _u32]
        L.ref_c3_mul.argtypes = [_u64p, _u64p, _u64p]
        L.ref_grinding.restype = _u64
        L.ref_grinding.argtypes = [_u64p, _u32]
        L.ref_num_threads.restype = ctypes.c_int

    @staticmethod
    def available():
        name = "libref_avx512.so" if _host_has_avx512() else "libref_avx2.so"
        return os.path.exists(os.path.join(REF_DIR, name))

    def lde(self, trace, n_ext):
        trace = np.ascontiguousarray(trace, dtype=np.uint64)
        n, ncols = trace.shape
        out = np.zeros((n_ext, ncols), dtype=np.uint64)
        self.L.ref_lde(_ptr(out), _ptr(trace), n_ext, n, ncols)
        return out % np.uint64(P) if out.size else out

    def ntt(self, a, inverse=False):
        """NTT_Goldilocks::NTT / INTT on a row-major (size, ncols) array."""
        a = np.array(a, dtype=np.uint64)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        self.L.ref_ntt(_ptr(a), a.shape[0], a.shape[1], int(inverse))
        return a % np.uint64(P)

    def compute_q(self, q, n, q_deg):
        """q: row-major (NExt, qDim) -> cmQ row-major (NExt, q_deg*qDim)."""
        q = np.array(q, dtype=np.uint64)
        n_ext, q_dim = q.shape
        out = np.zeros((n_ext, q_deg * q_dim), dtype=np.uint64)
        self.L.ref_compute_q(_ptr(out), _ptr(q), n, n_ext, q_deg, q_dim)
        return out % np.uint64(P)

    def merkletree(self, src, arity, tree_len):
        # The reference's batch leaf/node loops work 8 rows at a time and touch
        # up to 8 rows (and 8 digests) past the end for nrows < 8, so both
        # buffers get 8 rows of slack; only the first tree_len words are returned.
        src = np.ascontiguousarray(src, dtype=np.uint64)
        rows, cols = src.shape
        padded = np.zeros((rows + 8, cols), dtype=np.uint64)
        padded[:rows] = src
        tree = np.zeros(tree_len + 8 * 4 * 16, dtype=np.uint64)
        rc = self.L.ref_merkletree(_ptr(tree), _ptr(padded), cols, rows, arity)
        assert rc == 0
        return tree[:tree_len] % np.uint64(P)

    def permute(self, state):
        st = np.array(state, dtype=np.uint64)
        assert self.L.ref_permute(_ptr(st), st.size) == 0
        return st

    def linear_hash(self, vals, width):
        v = np.ascontiguousarray(vals, dtype=np.uint64)
        out = np.zeros(4, dtype=np.uint64)
        dummy = np.zeros(1, np.uint64)
        assert self.L.ref_linear_hash(_ptr(out), _ptr(v) if v.size else _ptr(dummy), v.size, width) == 0
        return out

    def c3_mul(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        r = np.zeros(3, dtype=np.uint64)
        self.L.ref_c3_mul(_ptr(r), _ptr(a), _ptr(b))
        return r

    def grinding(self, challenge, n_bits):
        ch = np.ascontiguousarray(np.asarray(challenge, dtype=np.uint64)[:3])
        return int(self.L.ref_grinding(_ptr(ch), n_bits))

    def num_threads(self):
        return self.L.ref_num_threads()
