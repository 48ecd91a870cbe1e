"""Helpers shared by the tests for comparing against the committed fixtures."""
import hashlib

import numpy as np


def lde_digest(lde_rowmajor):
    """sha256 of the canonical row-major LDE as little-endian u64 words."""
    return hashlib.sha256(np.ascontiguousarray(lde_rowmajor, dtype="<u8").tobytes()).hexdigest()


def colmajor_to_rowmajor(flat, n_rows, n_cols):
    return np.ascontiguousarray(np.asarray(flat, dtype=np.uint64).reshape(n_cols, n_rows).T)


def device_colmajor_sha256(pc, data, n_rows, n_cols, block_rows=1 << 20):
    """sha256 of a column-major device array (column c at data[c*n_rows ..]) as
    the row-major array it represents, transposed on the device in blocks of rows
    (so that the host never holds more than one block)."""
    h = hashlib.sha256()
    blk = min(block_rows, n_rows)
    out = pc.empty(blk * n_cols)
    for r0 in range(0, n_rows, blk):
        pc.transpose_cm2rm(out, data[r0:], blk, n_cols, n_rows, n_cols)
        h.update(np.ascontiguousarray(pc.to_host(out), dtype="<u8"))
    return h.hexdigest()
