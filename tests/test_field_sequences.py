"""Word-level models of the device's reduction sequences, checked against exact
integers on the CPU.

The device code forms x * 2^S (the in-register DFT twiddles w_16^k = 2^(12k) and
w_16^-k = -2^(96 - 12k)) as r = A*(2^32 - 1) + lo - (B1:B0) and then applies the
product reduction's correction: + (2^32 - 1) when the v_mad_u64_u32 carries out
and the subtraction does not borrow, - (2^32 - 1) in the opposite case. The model
below follows those instructions word by word (32-bit borrows and carries,
64-bit wrap), so a wrong term, a wrong shift or a correction that can wrap fails
here for some S or some input word.
"""
import random

import pytest

P = 2**64 - 2**32 + 1
M32, M64 = 2**32 - 1, 2**64 - 1


def pow2_terms(x, s):
    """(A, lo, B0, B1) of x * 2^s, as gl_pow2_terms builds them."""
    if s < 64:
        h = x >> (64 - s)
        return h & M32, (x << s) & M64, h >> 32, 0
    t = s - 64
    q = x >> (32 - t)
    return ((x & M32) << t) & M32, 0, q & M32, q >> 32


def pow2_reduce(a, lo, b0, b1):
    """gl_pow2_reduce1's instructions on 32/64-bit words."""
    r = a * M32 + lo                      # v_mad_u64_u32 r, c, A, -1, lo
    c, r = r >> 64, r & M64
    r0, r1 = r & M32, r >> 32
    d = r0 - b0                           # v_sub_co_u32
    b, r0 = int(d < 0), d & M32
    d = r1 - b1 - b                       # v_subb_co_u32
    b, r1 = int(d < 0), d & M32
    m, n = c & (1 - b), b & (1 - c)       # s_andn2_b64 x2
    tm, tn = (M32 if m else 0), (M32 if n else 0)  # v_cndmask_b32 x2
    s = r0 + tm + n                       # v_addc_co_u32 (carry-in n)
    r0, cy = s & M32, s >> 32
    r1 = (r1 + tn + cy) & M32             # v_addc_co_u32
    return (r1 << 32) | r0


EDGE = [0, 1, 2, M32, M32 + 1, P - 1, P, P + 1, M64, M64 - 1, 2**63, 2**63 - 1, M32 << 32, M32 << 16, (1 << 48) - 1]


def words(seed, n):
    rnd = random.Random(seed)
    out = list(EDGE)
    out += [rnd.getrandbits(64) for _ in range(n)]
    out += [rnd.getrandbits(rnd.randrange(1, 65)) for _ in range(n // 2)]
    out += [M64 - rnd.getrandbits(rnd.randrange(1, 40)) for _ in range(n // 4)]
    return out


@pytest.mark.parametrize("s", sorted({12 * k for k in range(1, 8)} | {96 - 12 * k for k in range(1, 8)}))
def test_dft_twiddle_shifts(s):
    """The shifts the DFT stages use, on many words (every result congruent and a 64-bit word)."""
    for x in words(s, 20000):
        got = pow2_reduce(*pow2_terms(x, s))
        assert 0 <= got <= M64
        assert got % P == x * pow(2, s, P) % P, (s, hex(x), hex(got))


def test_every_shift_below_96():
    """Every S in 1..95 (the bounds that keep the correction from wrapping hold for all of them)."""
    xs = words(7, 600)
    for s in range(1, 96):
        for x in xs:
            assert pow2_reduce(*pow2_terms(x, s)) % P == x * pow(2, s, P) % P, (s, hex(x))


def test_inverse_roots_are_negated_powers_of_two():
    """w_16 = 2^12 and w_16^-k = -2^(96 - 12k), the identity the DIF/DIT sign folding relies on."""
    w16 = 4096
    assert pow(w16, 16, P) == 1 and pow(w16, 8, P) == P - 1
    for k in range(1, 8):
        assert pow(w16, 16 - k, P) == (P - pow(2, 96 - 12 * k, P)) % P
