"""Seeded synthetic Goldilocks data (uniform in [0, p)).

Counter-based SplitMix64: element i = mix(seed + (i+1)*0x9E3779B97F4A7C15), re-mixed
while >= p. The GPU fill kernel (pil2hip_fill_uniform) computes the same values,
so large inputs are generated directly in HBM.
"""
import numpy as np

P = np.uint64(0xFFFFFFFF00000001)
_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_R = np.uint64(0xD1B54A32D192ED03)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(n, seed, start=0):
    """Elements [start, start+n) of the seeded stream as a uint64 array."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = _mix(np.uint64(seed) + i * _G)
        bad = z >= P
        while bad.any():
            z[bad] = _mix(z[bad] ^ _R)
            bad = z >= P
    return z


def challenges(n, seed):
    """n cubic-extension challenges (3 words each) from the seeded stream."""
    return uniform(3 * n, seed).reshape(n, 3)


def raw_words(n, seed):
    """n raw 64-bit SplitMix64 words (no reduction), for drawing shapes."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        return _mix(np.uint64(seed) + i * _G)


def batched_workload(seed=0x5EED0003, n_instances=64, min_log=12, max_log=21, min_cols=8, max_cols=256):
    """Shapes of the mixed-size batch: per instance log2(rows) uniform over
    [min_log, max_log], columns uniform over [min_cols, max_cols], blowup 2 or 4
    by a seeded coin, and a data seed for its uniform trace values.
    Returns a list of dicts {n_bits, n_bits_ext, n_cols, seed}."""
    w = raw_words(4 * n_instances, seed).reshape(n_instances, 4)
    out = []
    for i in range(n_instances):
        nb = min_log + int(w[i, 0] % np.uint64(max_log - min_log + 1))
        nc = min_cols + int(w[i, 1] % np.uint64(max_cols - min_cols + 1))
        blow = 1 if int(w[i, 2]) & 1 else 2
        out.append({"n_bits": nb, "n_bits_ext": nb + blow, "n_cols": nc, "seed": int(w[i, 3])})
    return out
