"""Helpers shared by the tests for comparing against the committed fixtures."""
import hashlib

import numpy as np


def lde_digest(lde_rowmajor):
    """sha256 of the canonical row-major LDE as little-endian u64 words."""
    return hashlib.sha256(np.ascontiguousarray(lde_rowmajor, dtype="<u8").tobytes()).hexdigest()


def colmajor_to_rowmajor(flat, n_rows, n_cols):
    return np.ascontiguousarray(np.asarray(flat, dtype=np.uint64).reshape(n_cols, n_rows).T)
