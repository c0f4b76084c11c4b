"""The binade-table identity behind the departure-form weight tables
(csrc/binade.hpp), checked on the CPU in float64.

For a line with displacement disp, element r departs from t = RN(r - disp).
Within one binade code of t (|t| in [2^b, 2^(b+1)), either sign) the rounding
grid is uniform and contains r, so alpha = t - floor(t) is the same for every
such r: one weight set per code equals the element-wise computation bit for
bit.  The kernels find each code's representative with a small candidate
search around ceil(2^b + disp) (positive codes) or floor(disp - 2^b)
(negative codes); both facts are checked here on random and adversarial
displacements (ties, exact binade edges, huge and tiny values).
"""
import math

import numpy as np
import pytest

K_BINADES = 16


def code_of(t: float):
    """binade_code(): e - 1 for t >= 1, -e for t <= -1 (|t| = m 2^e), else None."""
    if not math.isfinite(t):
        return None
    m, e = math.frexp(t)
    if t >= 1.0 and e <= K_BINADES:
        return e - 1
    if t <= -1.0 and e <= K_BINADES:
        return -e
    return None


def representative(disp: float, q: int):
    """The kernels' candidate search (binade_weights / binade_weights_neg):
    alpha of the first candidate whose departure point has code q, or None."""
    if q >= 0:
        lo = math.ldexp(1.0, q)
        hi = 2.0 * lo
        c0 = math.ceil(lo + disp)
        for d in (-1, 0, 1, 2):
            t = float(c0 + d) - disp
            if lo <= t < hi:
                return t - math.floor(t)
        return None
    b = -q - 1
    lo = math.ldexp(1.0, b)
    hi = 2.0 * lo
    c0 = math.floor(disp - lo)
    for d in (1, 0, -1, -2):
        t = float(c0 + d) - disp
        if lo <= -t < hi:
            return t - math.floor(t)
    return None


def _displacements():
    rng = np.random.default_rng(7)
    vals = list(rng.uniform(-40.0, 40.0, 60)) + list(rng.normal(size=20) * 1e-3)
    vals += [0.0, 0.5, -0.5, 1.5, 2.5, -2.5, 1e-300, -1e-300, 3e-17, 1.0 - 2.0 ** -53, 255.5, -1000.25,
             math.ldexp(1.0, -40) + 0.25, 0.1 * 3, 7.0 / 3.0, 4096.0 - 2.0 ** -40]
    return vals


@pytest.mark.parametrize("n", [300, 4200])
def test_alpha_depends_only_on_the_binade_code(n):
    for disp in _displacements():
        seen = {}
        for r in range(n):
            t = float(r) - disp
            q = code_of(t)
            if q is None:
                continue
            alpha = t - math.floor(t)
            if q in seen:
                assert seen[q] == alpha, (disp, r, q)
            else:
                seen[q] = alpha
        for q, alpha in seen.items():
            assert representative(disp, q) == alpha, (disp, q)


def test_representative_search_finds_every_reachable_code_on_long_lines():
    rng = np.random.default_rng(11)
    r = np.arange(0, 1 << 16, dtype=np.float64)
    for disp in rng.uniform(-3000.0, 3000.0, 25):
        t = r - disp
        for q in sorted({code_of(float(x)) for x in t[:: 97]} - {None}):
            assert representative(float(disp), q) is not None, (disp, q)


def test_codes_rise_with_the_departure_point():
    """segment_codes() takes the codes of a segment's end points as the range."""
    xs = sorted(np.random.default_rng(3).uniform(-70000, 70000, 4000))
    codes = [code_of(x) for x in xs]
    seq = [c for c in codes if c is not None]
    assert seq == sorted(seq)
