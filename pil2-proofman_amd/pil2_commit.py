"""Host-side mirror of pil2-stark's commitment interface over libpil2_commit_hip.so.

The classes keep the reference's names and argument meaning so code written
against the C++ API reads the same:

  NTT_Goldilocks.LDE         pil2-stark/src/goldilocks/src/ntt_goldilocks.cpp:369-404
  MerkleTreeGL               pil2-stark/src/starkpil/merkleTree/merkleTreeGL.{hpp,cpp}
                             (merkelize :316-336, getRoot :103-107, getGroupProof
                             :145-175, getLevel :84-101, getTreeNumElements
                             goldilocks_tooling.hpp:9-21)
  FRI.fold / FRI.merkelize   pil2-stark/src/starkpil/fri/fri.hpp:32-109
  write_custom_commit        pil2-stark/src/api/starks_api.hpp:96
  set_hash_family            pil2-stark/src/api/starks_api.hpp:28

Device buffers are torch tensors (dtype int64 holding the u64 bit pattern) on a
HIP device; torch is only the allocator and stream provider. Every operation
runs in the HIP library; if the library is missing this module raises at
import time — there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpil2_commit_hip.so")
P = 0xFFFFFFFF00000001
HASH_SIZE = 4

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libpil2_commit_hip.so is not built (run `python pil2-proofman_amd/build.py` "
        "or __graft_entry__.build()); no CPU fallback exists")

_lib = ctypes.CDLL(LIB_PATH)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64

_SIGS = {
    "set_hash_family": (None, [ctypes.c_uint8]),
    "write_custom_commit": (None, [_vp, _u64, _u64, _u64, _u64, _vp, _vp, ctypes.c_char_p]),
    "gen_device_buffers": (_vp, [ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32, ctypes.c_uint32]),
    "free_device_buffers": (None, [_vp]),
    "get_num_gpus": (_u64, []),
    "set_gpu_mode": (ctypes.c_bool, [ctypes.c_bool]),
    "pil2hip_tree_num_elements": (_u64, [_u64, _u64]),
    "pil2hip_lde": (_i64, [_vp, _vp, _vp, _u64, _u64, _u64, _vp]),
    "pil2hip_merkelize": (_i64, [_vp, _vp, _u64, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_commit": (_i64, [_vp, _vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_commit_noalloc": (_i64, [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_const_tree_size": (_u64, [_u64, _u64, _u64]),
    "pil2hip_const_tree": (_i64, [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_calculate_const_tree": (_i64, [_vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_write_const_tree": (_i64, [_vp, _u64, _u64, _u64, ctypes.c_char_p]),
    "pil2hip_load_const_tree": (_i64, [_vp, ctypes.c_char_p, _u64, ctypes.c_char_p]),
    "pil2hip_ntt": (_i64, [_vp, _vp, _vp, _u64, _u64, ctypes.c_bool, ctypes.c_bool, _vp]),
    "pil2hip_compute_q": (_i64, [_vp, _vp, _vp, _u64, _u64, _vp, _u64, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_commit_batch": (_i64, [_vp, _u64, ctypes.c_uint32, _vp, _vp]),
    "pil2hip_commit_batch_plan": (_i64, [_vp, _u64, _vp, _vp, _vp, _vp]),
    "pil2hip_batch_shape_plan": (_i64, [_u64, _u64, _u64, _u64, _vp]),
    "pil2hip_fri_layers": (_i64, [_vp, _vp, _vp, _u64, _vp, _u64, _u64, _vp, _vp, ctypes.c_bool, _vp]),
    "pil2hip_fri_release": (None, []),
    "pil2hip_batch_release": (None, []),
    "pil2hip_group_proofs": (_i64, [_vp, _vp, _u64, _u64, _u64, _vp, _u64, _u64, _u64, _vp, _u64, _vp]),
    "pil2hip_proof_levels": (_u64, [_u64, _u64, _u64]),
    "pil2hip_grinding": (_i64, [_vp, _vp, ctypes.c_uint32, _vp]),
    "pil2hip_transcript_create": (_vp, [_u64]),
    "pil2hip_transcript_put": (_i64, [_vp, _vp, _u64, _vp]),
    "pil2hip_transcript_put_host": (_i64, [_vp, _vp, _u64, _vp]),
    "pil2hip_transcript_get_field": (_i64, [_vp, _vp, _vp]),
    "pil2hip_transcript_get_state": (_i64, [_vp, _vp, _u64, _vp]),
    "pil2hip_transcript_get_permutations": (_i64, [_vp, _vp, _u64, _u64, _vp]),
    "pil2hip_transcript_destroy": (None, [_vp]),
    "pil2hip_proof_words": (_u64, [_u64, _u64, _u64, _u64]),
    "pil2hip_serve_create": (_vp, [_vp]),
    "pil2hip_serve_submit": (_i64, [_vp, _vp]),
    "pil2hip_serve_poll": (_i64, [_vp, _vp, _u64, ctypes.c_int32]),
    "pil2hip_serve_cancel": (_i64, [_vp, _u64]),
    "pil2hip_serve_open": (_i64, [_vp, _u64, _vp, _u64, _vp]),
    "pil2hip_serve_release": (_i64, [_vp, _u64]),
    "pil2hip_serve_stats": (_i64, [_vp, _vp]),
    "pil2hip_serve_now_us": (ctypes.c_double, [_vp]),
    "pil2hip_serve_job_bytes": (_u64, [_u64, _u64, _u64, _u64, _u64, _u64, ctypes.c_bool]),
    "pil2hip_serve_destroy": (None, [_vp]),
    "pil2hip_get_root": (_i64, [_vp, _vp, _u64, _u64, _vp]),
    "pil2hip_level_words": (_u64, [_u64, _u64]),
    "pil2hip_get_level": (_i64, [_vp, _vp, _u64, _u64, _u64, _vp]),
    "pil2hip_commit_witness": (_i64, [_vp, _u64, _u64, _u64, _u64, _vp, _vp, _u64, _u64, _vp]),
    "pil2hip_fri_fold": (_i64, [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_fri_transpose": (_i64, [_vp, _vp, _u64, _u64, _vp]),
    "pil2hip_transpose_cm2rm": (_i64, [_vp, _vp, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_p2_permute": (_i64, [_vp, _u64, ctypes.c_uint32, _vp]),
    "pil2hip_fill_uniform": (_i64, [_vp, _u64, _u64, _u64, _vp]),
    "pil2hip_mul_peak": (_i64, [_vp, _vp, _u64, ctypes.c_uint32, _vp]),
    "pil2hip_commit_host": (_i64, [_vp, _vp, _vp, _vp, _u64, _u64, _u64, _u64, _u64, _vp]),
    "pil2hip_commit_host_stage_words": (_u64, [_u64, _u64, _u64, _u64]),
    "pil2hip_gl_binop": (_i64, [_vp, _vp, _vp, _u64, ctypes.c_int32, _vp]),
    "pil2hip_timing_enable": (None, [ctypes.c_bool]),
    "pil2hip_timing_read": (_u64, [_vp, ctypes.c_bool]),
    "pil2hip_last_error": (ctypes.c_char_p, []),
    "pil2hip_set_abort_on_error": (None, [ctypes.c_bool]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

lib = _lib


class HipError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise HipError("libpil2_commit_hip error %d: %s" % (rc, _lib.pil2hip_last_error().decode()))


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def to_device(arr, device="cuda"):
    """numpy uint64 array -> int64 device tensor with the same bits."""
    torch = _torch()
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64)).to(device)


def to_host(t):
    """device tensor (int64 bits) -> numpy uint64 array."""
    return t.detach().cpu().numpy().view(np.uint64)


def empty(n, device="cuda", stream=None):
    """Device int64 tensor of n words. With `stream` (a torch stream) the block is
    allocated in that stream's context, so the caching allocator hands it out again
    only after the work queued on `stream` has used it: temporaries that this module
    allocates for an asynchronous call on a caller's stream stay valid until the
    call's kernels have run."""
    torch = _torch()
    if stream is None:
        return torch.empty(int(n), dtype=torch.int64, device=device)
    with torch.cuda.stream(stream):
        return torch.empty(int(n), dtype=torch.int64, device=device)


# ------------------------------------------------------------------ ABI --
def set_hash_family(fam):
    """starks_api.hpp:28 — 1 = Poseidon1 (rejected by this library), 2 = Poseidon2."""
    _lib.set_hash_family(int(fam))


def set_abort_on_error(flag):
    _lib.pil2hip_set_abort_on_error(bool(flag))


def get_tree_num_elements(degree, arity=2):
    """goldilocks_tooling.hpp:9-21"""
    return int(_lib.pil2hip_tree_num_elements(int(degree), int(arity)))


def write_custom_commit(root, arity, n_bits, n_bits_ext, n_cols, d_buffers, buffer, buffer_file=""):
    """starks_api.hpp:96. root: numpy uint64[4] (out); buffer: numpy row-major trace."""
    assert root.dtype == np.uint64 and root.size >= 4
    buf = np.ascontiguousarray(buffer, dtype=np.uint64)
    _lib.write_custom_commit(root.ctypes.data_as(_vp), arity, n_bits, n_bits_ext, n_cols,
                             d_buffers, buf.ctypes.data_as(_vp), buffer_file.encode())
    err = _lib.pil2hip_last_error().decode()
    if err:
        raise HipError("write_custom_commit: " + err)


def commit_witness(root, arity, n_bits, n_bits_ext, n_cols, trace, aux_trace, offset_dst, offset_mt, d_buffers=None):
    """commit_witness_cpu's data path (starks_api.cpp:683-723) through
    pil2hip_commit_witness: trace (host uint64, row-major 2^n_bits x n_cols) ->
    aux_trace[offset_dst:] = row-major LDE, aux_trace[offset_mt:] = Merkle nodes,
    root (4 words). aux_trace and root are host uint64 arrays written in place."""
    tr = np.ascontiguousarray(trace, dtype=np.uint64)
    for a in (root, aux_trace):
        if not (isinstance(a, np.ndarray) and a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]):
            raise ValueError("commit_witness: root and aux_trace must be contiguous uint64 arrays")
    ne = 1 << n_bits_ext
    if root.size < 4 or aux_trace.size < max(offset_dst + ne * n_cols, offset_mt + get_tree_num_elements(ne, arity)):
        raise ValueError("commit_witness: root or aux_trace too small")
    _check(_lib.pil2hip_commit_witness(root.ctypes.data_as(_vp), arity, n_bits, n_bits_ext, n_cols,
                                       tr.ctypes.data_as(_vp), aux_trace.ctypes.data_as(_vp), offset_dst, offset_mt,
                                       d_buffers))


def const_tree_size(n_bits_ext, n_constants, arity):
    """ConstTree::getConstTreeSizeGL (const_pols.hpp:25-30), in words."""
    return int(_lib.pil2hip_const_tree_size(n_bits_ext, n_constants, arity))


def calculate_const_tree(const_pols, n_bits, n_bits_ext, arity, d_buffers=None):
    """ConstTree::calculateConstTreeGL (const_pols.hpp:33-44): host row-major
    2^n_bits x nConstants in; the host const tree buffer (row-major LDE, then the
    Merkle nodes) out."""
    pols = np.ascontiguousarray(const_pols, dtype=np.uint64)
    n_const = pols.shape[1] if pols.ndim == 2 else 0
    tree = np.zeros(const_tree_size(n_bits_ext, n_const, arity), dtype=np.uint64)
    _check(_lib.pil2hip_calculate_const_tree(pols.ctypes.data_as(_vp) if pols.size else None,
                                             tree.ctypes.data_as(_vp), n_bits, n_bits_ext, n_const, arity, d_buffers))
    return tree


def write_const_tree(tree, n_bits_ext, n_constants, arity, path):
    """write_const_tree / MerkleTreeGL::writeFile (merkleTreeGL.cpp:338-345)."""
    tree = np.ascontiguousarray(tree, dtype=np.uint64)
    _check(_lib.pil2hip_write_const_tree(tree.ctypes.data_as(_vp), n_bits_ext, n_constants, arity,
                                         os.fsencode(path)))


def load_const_tree(path, size, verkey_path=None):
    """ConstTree::loadConstTree, GL branch (const_pols.hpp:66-103): returns
    (loaded-and-verified, buffer)."""
    tree = np.zeros(int(size), dtype=np.uint64)
    ok = _lib.pil2hip_load_const_tree(tree.ctypes.data_as(_vp), os.fsencode(path), int(size),
                                      os.fsencode(verkey_path) if verkey_path else None)
    if ok < 0:
        raise HipError("load_const_tree: " + _lib.pil2hip_last_error().decode())
    return bool(ok), tree


def gen_device_buffers(arity=3, max_n_bits_ext=24):
    return _lib.gen_device_buffers(0, 1, None, arity, max_n_bits_ext)


def free_device_buffers(handle):
    _lib.free_device_buffers(handle)


# ----------------------------------------------------------------- NTT --
class NTT_Goldilocks:
    """Mirror of NTT_Goldilocks (ntt_goldilocks.hpp:66-193): NTT, INTT and LDE on device."""

    def __init__(self, max_domain_size):
        self.n = int(max_domain_size)

    @staticmethod
    def _log2(n):
        lg = int(n).bit_length() - 1
        if (1 << lg) != n:
            raise ValueError("size must be a power of two")
        return lg

    def NTT(self, dst, src, size, ncols=1, buffer=None, inverse=False, row_major=True, stream=None):
        """ntt_goldilocks.cpp:211-260 on device tensors: natural order in and out;
        row-major size x ncols (the CPU layout) or column-major (row_major=False).
        dst None (or src) transforms in place. buffer: device scratch, allocated
        when needed (2*size*ncols words row-major, size*ncols in-place column-major)."""
        if size == 0 or ncols == 0:
            return
        lg = self._log2(size)
        dst = src if dst is None else dst
        need = 0 if lg < 4 else (2 * size * ncols if row_major else (size * ncols if dst.data_ptr() == src.data_ptr()
                                                                       else 0))
        if need and (buffer is None or buffer.numel() < need):
            buffer = empty(need, device=src.device, stream=stream)
        _check(_lib.pil2hip_ntt(_ptr(dst), _ptr(src), _ptr(buffer) if need else None, lg, ncols, bool(inverse),
                                bool(row_major), _stream(stream)))

    def INTT(self, dst, src, size, ncols=1, buffer=None, row_major=True, stream=None):
        """ntt_goldilocks.hpp:187-191: the inverse transform (includes 1/size)."""
        self.NTT(dst, src, size, ncols, buffer, inverse=True, row_major=row_major, stream=stream)

    def LDE(self, output, input, n_extended, n, ncols, buffer=None, stream=None):
        """ntt_goldilocks.cpp:369-404. input: device row-major n x ncols; output:
        device column-major ncols x n_extended; buffer: device scratch >= n*ncols."""
        if n == 0 or ncols == 0:
            return
        if buffer is None:
            buffer = empty(n * ncols, device=input.device, stream=stream)
        _check(_lib.pil2hip_lde(_ptr(output), _ptr(input), _ptr(buffer), self._log2(n),
                                self._log2(n_extended), ncols, _stream(stream)))


# ---------------------------------------------------------------- tree --
class MerkleTreeGL:
    """Mirror of MerkleTreeGL over device buffers (merkleTreeGL.hpp:12-115).

    source: device tensor; element (row, col) at source[row*row_stride + col*col_stride]
    (column-major LDE: row_stride=1, col_stride=height; row-major: width, 1).
    """

    def __init__(self, arity, last_level_verification, custom, height, width,
                 source=None, row_stride=None, col_stride=None, nodes=None):
        self.arity = int(arity)
        self.last_level_verification = int(last_level_verification)
        self.custom = custom
        self.height = int(height)
        self.width = int(width)
        self.num_nodes = get_tree_num_elements(self.height, self.arity)
        self.source = source
        self.row_stride = self.width if row_stride is None else int(row_stride)
        self.col_stride = 1 if col_stride is None else int(col_stride)
        self.nodes = nodes

    def set_source(self, source, row_stride=None, col_stride=None):
        self.source = source
        if row_stride is not None:
            self.row_stride, self.col_stride = int(row_stride), int(col_stride)

    def set_nodes(self, nodes):
        self.nodes = nodes

    def merkelize(self, stream=None):
        """merkleTreeGL.cpp:316-336"""
        if self.nodes is None:
            self.nodes = empty(self.num_nodes, device=self.source.device, stream=stream)
        _check(_lib.pil2hip_merkelize(_ptr(self.nodes), _ptr(self.source), self.height, self.width,
                                      self.row_stride, self.col_stride, self.arity, _stream(stream)))

    def get_root(self, stream=None):
        """merkleTreeGL.cpp:103-107"""
        root = np.zeros(4, dtype=np.uint64)
        _check(_lib.pil2hip_get_root(root.ctypes.data_as(_vp), _ptr(self.nodes), self.height,
                                     self.arity, _stream(stream)))
        return root

    def get_merkle_proof_length(self):
        """merkleTreeGL.cpp:50-55"""
        if self.height > 1:
            import math
            return int(math.ceil(math.log2(self.height) / math.log2(self.arity))) - self.last_level_verification
        return 0

    def get_level(self, stream=None):
        """merkleTreeGL.cpp:84-101 (pil2hip_get_level): the nodes of the level with
        <= arity^llv entries, zero-padded to arity^llv digests; empty for llv 0."""
        llv = self.last_level_verification
        out = np.zeros(int(_lib.pil2hip_level_words(self.arity, llv)), dtype=np.uint64)
        if out.size:
            _check(_lib.pil2hip_get_level(out.ctypes.data_as(_vp), _ptr(self.nodes), self.height, self.arity, llv,
                                          _stream(stream)))
        return out

    def get_group_proof(self, idx):
        """merkleTreeGL.cpp:145-175: row values then sibling digests per level."""
        torch = _torch()
        row_idx = torch.arange(self.width, device=self.source.device, dtype=torch.int64) * self.col_stride
        vals = to_host(self.source.reshape(-1)[idx * self.row_stride + row_idx])
        nodes = to_host(self.nodes)
        out = [vals]
        n, offset, llv, a = self.height, 0, self.last_level_verification, self.arity
        lim = a ** llv
        while (n > 1) if llv == 0 else (n > lim):
            cur = idx % a
            si = idx - cur
            for i in range(a):
                if i != cur:
                    out.append(nodes[(offset + si + i) * 4:(offset + si + i) * 4 + 4])
            nxt = (n + a - 1) // a
            offset += nxt * a
            idx //= a
            n = nxt
        return np.concatenate(out)


# ----------------------------------------------------------------- FRI --
class FRI:
    """Mirror of FRI<Goldilocks::Element> (fri.hpp) fold and per-layer commit."""

    @staticmethod
    def steps(n_bits_ext, folding=3, final_degree=5):
        """FRI step schedule (setup/pil2-stark/src/types/stark_struct.rs:162-166):
        nBits of every layer, from the LDE domain down to the final polynomial."""
        out, b = [n_bits_ext], n_bits_ext
        while b > final_degree + 1:
            b = b - folding if b > folding + final_degree else final_degree
            out.append(b)
        return out

    @staticmethod
    def fold(step, pol, challenge, n_bits_ext, prev_bits, current_bits, out=None, stream=None):
        """fri.hpp:32-99. pol: device tensor of 2^prev_bits cubic values (3 words each).
        Folds in place unless `out` is given; returns the tensor holding the result."""
        ch = np.ascontiguousarray(challenge, dtype=np.uint64)
        dst = pol if out is None else out
        _check(_lib.pil2hip_fri_fold(_ptr(dst), _ptr(pol), ch.ctypes.data_as(_vp), step, n_bits_ext,
                                     prev_bits, current_bits, _stream(stream)))
        return dst

    @staticmethod
    def layers(pol, work, challenges, n_bits_ext, steps, trees, arity=3, graph=True, stream=None):
        """The FRI layer loop (pil2hip_fri_layers): fold step s > 0 with the device
        challenges[3s:3s+3] into `work`, commit layer s < len(steps)-1 into
        trees[s] (MerkleTreeGL of 2^steps[s+1] rows; its source and nodes tensors
        are allocated here when absent). graph=True replays one cached hipGraph."""
        n = len(steps)
        for s_ in range(n - 1):
            t = trees[s_]
            if t.source is None:
                t.source = empty(3 << steps[s_], device=pol.device, stream=stream)
            if t.nodes is None:
                t.nodes = empty(t.num_nodes, device=pol.device, stream=stream)
            t.row_stride, t.col_stride = t.width, 1
        st = (ctypes.c_uint64 * n)(*steps)
        srcs = (_vp * max(n - 1, 1))(*[_ptr(t.source) for t in trees[:n - 1]])
        nds = (_vp * max(n - 1, 1))(*[_ptr(t.nodes) for t in trees[:n - 1]])
        _check(_lib.pil2hip_fri_layers(_ptr(pol), _ptr(work), _ptr(challenges), n_bits_ext, st, n, arity,
                                       srcs, nds, graph, _stream(stream)))

    @staticmethod
    def prove_queries(tree, queries, current_bits, stream=None):
        """FRI::proveFRIQueries (fri.hpp:131-139): group proofs on a layer tree at
        the query indices reduced mod 2^current_bits. Returns (n, proof_words) uint64."""
        q = np.asarray(queries, dtype=np.uint64) % np.uint64(1 << current_bits)
        pw = proof_words(tree.height, tree.width, tree.arity, tree.last_level_verification)
        out = group_proofs(tree.source, tree.height, tree.width, tree.nodes, tree.arity,
                           tree.last_level_verification, to_device(q, tree.source.device),
                           row_stride=tree.row_stride, col_stride=tree.col_stride, stream=stream)
        return to_host(out)[:q.size * pw].reshape(q.size, pw)

    @staticmethod
    def merkelize(tree, pol, current_bits, next_bits, stream=None):
        """fri.hpp:102-109: regroup (getTransposed) into tree.source, merkelize, root."""
        degree = 1 << current_bits
        if tree.source is None:
            tree.source = empty(degree * 3, device=pol.device, stream=stream)
        _check(_lib.pil2hip_fri_transpose(_ptr(tree.source), _ptr(pol), degree, next_bits, _stream(stream)))
        tree.row_stride, tree.col_stride = tree.width, 1
        tree.merkelize(stream)
        return tree.get_root(stream)


def p2_permute(states, width, stream=None):
    """Poseidon2Goldilocks<width>::permute over consecutive states (device tensor)."""
    count = states.numel() // width
    _check(_lib.pil2hip_p2_permute(_ptr(states), count, width, _stream(stream)))


def fill_uniform(out, seed, start=0, stream=None):
    """Fill a device tensor with the seeded uniform stream (datagen.uniform on device)."""
    _check(_lib.pil2hip_fill_uniform(_ptr(out), out.numel(), seed, start, _stream(stream)))
    return out


MUL_PEAK_CHAINS = 18


def commit_host(h_trace, arity, n_bits, n_bits_ext, n_cols, lde=None, nodes=None, stage=None, group_cols=0,
                stream=None):
    """pil2hip_commit for a host trace in page-locked memory (a contiguous pinned
    CPU tensor), with the column groups' transfer overlapped. Returns (lde, nodes)."""
    torch = _torch()
    ne = 1 << n_bits_ext
    if not torch.is_tensor(h_trace) or h_trace.is_cuda or not h_trace.is_contiguous():
        raise ValueError("commit_host: h_trace must be a contiguous pinned host tensor")
    ptr, numel = h_trace.data_ptr(), h_trace.numel()
    if numel < (1 << n_bits) * n_cols:
        raise ValueError("commit_host: trace smaller than 2^n_bits * n_cols")
    dev = torch.device("cuda", torch.cuda.current_device())
    lde = empty(max(ne * n_cols, 1), dev, stream) if lde is None else lde
    nodes = empty(get_tree_num_elements(ne, arity), dev, stream) if nodes is None else nodes
    words = _lib.pil2hip_commit_host_stage_words(n_bits, n_cols, arity, group_cols)
    stage = empty(max(words, 1), dev, stream) if stage is None else stage
    if stage.numel() < words:
        raise ValueError("commit_host: stage smaller than %d words" % words)
    _check(_lib.pil2hip_commit_host(_ptr(lde), _ptr(nodes), ctypes.c_void_p(ptr), _ptr(stage), arity, n_bits,
                                    n_bits_ext, n_cols, group_cols, _stream(stream)))
    return lde, nodes


def mul_peak(out, inp, lanes, iters, stream=None):
    """Product throughput calibration: out[18i+k] = inp[36i+k] * inp[36i+18+k]^iters
    mod p, as lanes*iters*18 chained products of the Poseidon2 kernels' multiply."""
    c = MUL_PEAK_CHAINS
    if inp.numel() < 2 * c * lanes or out.numel() < c * lanes:
        raise ValueError("mul_peak: buffers smaller than 36*lanes / 18*lanes words")
    _check(_lib.pil2hip_mul_peak(_ptr(out), _ptr(inp), lanes, iters, _stream(stream)))
    return out


def transpose_cm2rm(out, inp, rows, cols, in_col_stride, out_row_stride, stream=None):
    _check(_lib.pil2hip_transpose_cm2rm(_ptr(out), _ptr(inp), rows, cols, in_col_stride,
                                        out_row_stride, _stream(stream)))


def commit(trace_dev, arity, n_bits, n_bits_ext, n_cols, lde=None, nodes=None, scratch=None, stream=None):
    """LDE + merkelize on device (commit_witness_cpu core, starks_api.cpp:683-723).
    Returns (lde column-major, nodes)."""
    n, ne = 1 << n_bits, 1 << n_bits_ext
    dev = trace_dev.device
    lde = empty(ne * n_cols, dev, stream) if lde is None else lde
    nodes = empty(get_tree_num_elements(ne, arity), dev, stream) if nodes is None else nodes
    scratch = empty(max(n * n_cols, 1), dev, stream) if scratch is None else scratch
    _check(_lib.pil2hip_commit(_ptr(lde), _ptr(nodes), _ptr(trace_dev), _ptr(scratch), arity, n_bits,
                               n_bits_ext, n_cols, _stream(stream)))
    return lde, nodes


def compute_q(q_dev, n_bits, n_bits_ext, q_deg, q_dim=3, arity=3, row_major=True, cmq=None, nodes=None,
              scratch=None, merkelize=True, stream=None):
    """Starks::computeQ (starks.hpp:187-239) on device: q_dev holds the quotient's
    NExt x q_dim extended-coset evaluations (row-major, or column-major with
    row_major=False). Returns (cmQ column-major q_deg*q_dim x NExt, nodes or None)."""
    n, ne = 1 << n_bits, 1 << n_bits_ext
    dev = q_dev.device
    cmq = empty(ne * q_deg * q_dim, dev, stream) if cmq is None else cmq
    if merkelize and nodes is None:
        nodes = empty(get_tree_num_elements(ne, arity), dev, stream)
    scratch = empty(max(n * q_deg * q_dim, 1), dev, stream) if scratch is None else scratch
    rs, cs = (q_dim, 1) if row_major else (1, ne)
    _check(_lib.pil2hip_compute_q(_ptr(cmq), _ptr(nodes) if merkelize else None, _ptr(q_dev), rs, cs,
                                  _ptr(scratch), n_bits, n_bits_ext, q_deg, q_dim, arity, _stream(stream)))
    return cmq, (nodes if merkelize else None)


class Pil2CommitJob(ctypes.Structure):
    """Mirror of the C struct Pil2CommitJob (include/pil2_commit_hip.h)."""
    _fields_ = [("d_trace", _vp), ("d_lde", _vp), ("d_nodes", _vp), ("d_scratch", _vp),
                ("n_bits", _u64), ("n_bits_ext", _u64), ("n_cols", _u64), ("arity", _u64)]


class Pil2CommitPlan(ctypes.Structure):
    """Mirror of the C struct Pil2CommitPlan (include/pil2_commit_hip.h); 0 = tune."""
    _fields_ = [("n_streams", ctypes.c_uint32), ("ntt_tile_log", ctypes.c_uint32),
                ("ntt_radix_log", ctypes.c_uint32), ("coop_max_parents", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class CommitBatch:
    """Many independent commits launched as one hipGraph (pil2hip_commit_batch).

    `jobs` is a list of dicts with keys trace (device tensor, row-major
    2^n_bits x n_cols), n_bits, n_bits_ext, n_cols and optionally lde, nodes,
    scratch (allocated here when absent). The tensors are kept alive by this
    object because the library caches the captured graph by their addresses."""

    def __init__(self, jobs, arity=3, device=None):
        self.jobs = []
        arr = (Pil2CommitJob * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            dev = j["trace"].device if device is None else device
            n, ne, nc = 1 << j["n_bits"], 1 << j["n_bits_ext"], j["n_cols"]
            lde = j.get("lde")
            lde = empty(ne * nc, dev) if lde is None else lde
            nodes = j.get("nodes")
            nodes = empty(get_tree_num_elements(ne, arity), dev) if nodes is None else nodes
            scratch = j.get("scratch")
            scratch = empty(max(n * nc, 1), dev) if scratch is None else scratch
            rec = dict(j, lde=lde, nodes=nodes, scratch=scratch, arity=arity)
            self.jobs.append(rec)
            arr[i] = Pil2CommitJob(j["trace"].data_ptr(), lde.data_ptr(), nodes.data_ptr(), scratch.data_ptr(),
                                   j["n_bits"], j["n_bits_ext"], nc, arity)
        self._arr = arr
        dev = jobs[0]["trace"].device if jobs else "cuda"
        self.roots = empty(4 * max(len(jobs), 1), dev)
        self.streams = None
        self.plan = None

    def launch(self, n_streams=0, stream=None, plan=None):
        """Enqueue the batch; returns the number of graph branches used. `plan`
        (a dict of Pil2CommitPlan fields: all missing or 0 = tune everything per
        shape; otherwise n_streams 0 = tune the branch count and other fields 0 =
        defaults) selects pil2hip_commit_batch_plan; the plan used is kept in
        self.plan."""
        if plan is None:
            rc = _lib.pil2hip_commit_batch(ctypes.cast(self._arr, _vp), len(self.jobs), n_streams,
                                           _ptr(self.roots), _stream(stream))
        else:
            want = Pil2CommitPlan(**{k: int(v) for k, v in plan.items()})
            got = Pil2CommitPlan()
            rc = _lib.pil2hip_commit_batch_plan(ctypes.cast(self._arr, _vp), len(self.jobs), ctypes.byref(want),
                                                ctypes.byref(got), _ptr(self.roots), _stream(stream))
            if rc >= 0:
                self.plan = got.as_dict()
        if rc < 0:
            _check(rc)
        self.streams = int(rc)
        return self.streams

    def roots_host(self):
        return to_host(self.roots)[:4 * len(self.jobs)].reshape(-1, 4)


def batch_shape_plan(n_bits, n_bits_ext, n_cols, arity):
    """(measured, plan dict) the tuned batch path uses for a job of this shape."""
    out = Pil2CommitPlan()
    rc = _lib.pil2hip_batch_shape_plan(n_bits, n_bits_ext, n_cols, arity, ctypes.byref(out))
    if rc < 0:
        _check(rc)
    return bool(rc), out.as_dict()


def batch_release():
    _lib.pil2hip_batch_release()


def proof_words(n_rows, n_cols, arity, llv):
    return int(_lib.pil2hip_proof_words(n_rows, n_cols, arity, llv))


def group_proofs(src, n_rows, n_cols, nodes, arity, llv, idx, row_stride=1, col_stride=None, stream=None):
    """Device getGroupProof for many row indices (idx: device int64 tensor).
    The source defaults to the column-major LDE layout (row_stride 1, col_stride
    n_rows). Returns a device tensor of len(idx) * proof_words(...) words."""
    pw = proof_words(n_rows, n_cols, arity, llv)
    cs = n_rows if col_stride is None else col_stride
    out = empty(max(idx.numel() * pw, 1), src.device, stream)
    _check(_lib.pil2hip_group_proofs(_ptr(out), _ptr(src), row_stride, cs, n_cols, _ptr(nodes), n_rows, arity, llv,
                                     _ptr(idx), idx.numel(), _stream(stream)))
    return out


# -------------------------------------------------------------- serving --
class Pil2ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _u64), ("n_streams", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32),
                ("arity", _u64), ("llv", _u64)]


class Pil2ServeRequest(ctypes.Structure):
    _fields_ = [("n_bits", _u64), ("n_bits_ext", _u64), ("n_cols", _u64), ("h_trace", _vp), ("d_trace", _vp),
                ("seed", _u64), ("n_queries", _u64), ("h_queries", _vp), ("h_root", _vp), ("h_proofs", _vp),
                ("tag", _u64), ("keep_resident", ctypes.c_uint32)]


class Pil2ServeCompletion(ctypes.Structure):
    _fields_ = [("job_id", _u64), ("tag", _u64), ("status", _i64), ("t_submit_us", ctypes.c_double),
                ("t_admit_us", ctypes.c_double), ("t_done_us", ctypes.c_double), ("stream", ctypes.c_uint32)]


class Pil2ServeStats(ctypes.Structure):
    _fields_ = [(k, _u64) for k in ("submitted", "admitted", "completed", "failed", "cancelled", "queued",
                                    "running", "resident_jobs", "bytes_in_use", "peak_bytes", "budget_bytes")] + \
               [("peak_running", ctypes.c_uint32)]


class ServeEngine:
    """Commit server (pil2hip_serve_*): jobs are admitted by HBM budget as they
    arrive and run concurrently; each completion carries the job's root and
    group proofs (host numpy arrays owned by the job record)."""

    def __init__(self, hbm_budget_bytes, n_streams=4, max_inflight=16, arity=3, llv=2):
        cfg = Pil2ServeConfig(int(hbm_budget_bytes), n_streams, max_inflight, arity, llv)
        self.arity, self.llv = arity, llv
        self.h = _lib.pil2hip_serve_create(ctypes.byref(cfg))
        if not self.h:
            raise HipError("pil2hip_serve_create: " + _lib.pil2hip_last_error().decode())
        self.records = {}

    def submit(self, n_bits, n_bits_ext, n_cols, queries=(), seed=0, h_trace=None, d_trace=None, tag=0,
               keep_resident=False):
        """Queue a job; returns its id. h_trace: numpy uint64 (rows, cols); d_trace: device tensor."""
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.uint64))
        pw = proof_words(1 << n_bits_ext, n_cols, self.arity, self.llv)
        rec = {"root": np.zeros(4, dtype=np.uint64), "proofs": np.zeros(max(q.size * pw, 1), dtype=np.uint64),
               "queries": q, "pw": pw, "n_bits": n_bits, "n_bits_ext": n_bits_ext, "n_cols": n_cols}
        if h_trace is not None:
            rec["h_trace"] = np.ascontiguousarray(h_trace, dtype=np.uint64)
        rec["d_trace"] = d_trace
        req = Pil2ServeRequest(n_bits, n_bits_ext, n_cols,
                               rec["h_trace"].ctypes.data if h_trace is not None else None,
                               d_trace.data_ptr() if d_trace is not None else None,
                               seed, q.size, q.ctypes.data if q.size else None, rec["root"].ctypes.data,
                               rec["proofs"].ctypes.data if q.size else None, tag, 1 if keep_resident else 0)
        jid = _lib.pil2hip_serve_submit(self.h, ctypes.byref(req))
        if jid < 0:
            _check(jid)
        self.records[jid] = rec
        return jid

    def poll(self, max_n=256, timeout_ms=100):
        buf = (Pil2ServeCompletion * max_n)()
        n = _lib.pil2hip_serve_poll(self.h, ctypes.cast(buf, _vp), max_n, timeout_ms)
        if n < 0:
            _check(n)
        out = []
        for c in buf[:n]:
            rec = self.records.get(c.job_id, {})
            rec.update(status=c.status, t_submit_us=c.t_submit_us, t_admit_us=c.t_admit_us, t_done_us=c.t_done_us,
                       stream=c.stream, tag=c.tag, job_id=c.job_id)
            out.append(rec)
        return out

    def cancel(self, job_id):
        return _lib.pil2hip_serve_cancel(self.h, job_id)

    def open(self, job_id, queries):
        rec = self.records[job_id]
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.uint64))
        out = np.zeros(max(q.size * rec["pw"], 1), dtype=np.uint64)
        _check(_lib.pil2hip_serve_open(self.h, job_id, q.ctypes.data_as(_vp), q.size, out.ctypes.data_as(_vp)))
        return out[:q.size * rec["pw"]]

    def release(self, job_id):
        _check(_lib.pil2hip_serve_release(self.h, job_id))

    def forget(self, job_id):
        """Release the job's resources in the engine and drop the host record."""
        rc = _lib.pil2hip_serve_release(self.h, job_id)
        self.records.pop(job_id, None)
        return rc

    def stats(self):
        st = Pil2ServeStats()
        _check(_lib.pil2hip_serve_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in Pil2ServeStats._fields_}

    def now_us(self):
        return _lib.pil2hip_serve_now_us(self.h)

    def close(self):
        if self.h:
            _lib.pil2hip_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_job_bytes(n_bits, n_bits_ext, n_cols, arity, n_queries, llv, device_trace=False):
    return int(_lib.pil2hip_serve_job_bytes(n_bits, n_bits_ext, n_cols, arity, n_queries, llv, device_trace))


def grinding(challenge, n_bits, stream=None):
    """Poseidon2Goldilocks<8>::grinding (poseidon2_goldilocks.cpp:177-232): the
    smallest nonce whose permuted state word 0 is below 2^(64 - n_bits)."""
    ch = np.ascontiguousarray(np.asarray(challenge, dtype=np.uint64)[:3])
    out = np.zeros(1, dtype=np.uint64)
    _check(_lib.pil2hip_grinding(out.ctypes.data_as(_vp), ch.ctypes.data_as(_vp), n_bits, _stream(stream)))
    return int(out[0])


class TranscriptGL:
    """Mirror of TranscriptGL (transcriptGL.hpp:11-58) with the sponge state in
    device memory (pil2hip_transcript_*)."""

    def __init__(self, arity, custom=False):
        self.h = _lib.pil2hip_transcript_create(arity)
        if not self.h:
            raise HipError("pil2hip_transcript_create: " + _lib.pil2hip_last_error().decode())

    def put(self, values, stream=None):
        """values: numpy/list (host) or a device tensor."""
        if hasattr(values, "data_ptr"):
            _check(_lib.pil2hip_transcript_put(self.h, _ptr(values), values.numel(), _stream(stream)))
        else:
            v = np.ascontiguousarray(np.asarray(values, dtype=np.uint64))
            if v.size:
                _check(_lib.pil2hip_transcript_put_host(self.h, v.ctypes.data_as(_vp), v.size, _stream(stream)))

    def get_field(self, stream=None):
        out = np.zeros(3, dtype=np.uint64)
        _check(_lib.pil2hip_transcript_get_field(self.h, out.ctypes.data_as(_vp), _stream(stream)))
        return out

    def get_state(self, n=4, stream=None):
        out = np.zeros(n, dtype=np.uint64)
        _check(_lib.pil2hip_transcript_get_state(self.h, out.ctypes.data_as(_vp), n, _stream(stream)))
        return out

    def get_permutations(self, n, n_bits, stream=None):
        out = np.zeros(max(n, 1), dtype=np.uint64)
        _check(_lib.pil2hip_transcript_get_permutations(self.h, out.ctypes.data_as(_vp), n, n_bits,
                                                        _stream(stream)))
        return out[:n]

    def close(self):
        if self.h:
            _lib.pil2hip_transcript_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
