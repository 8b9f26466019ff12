"""The resampler's reciprocal: one fused Newton step from a faithful approximation.

The resampler projects a sample with u = fma(x kx, RN(1 / y), 8), the rounding of the
reference's `1.0f / y` (IMRTOpt/src/IMRTgeom.cu:370-516).  On the device RN(1 / y) is formed from
the hardware reciprocal r (within one ulp, so one of the two floats around 1 / y) as
fma(r, fma(-y, r, 1), r).  This checks, over every significand of [1, 2) and for both
candidates r, that the step returns the correctly rounded quotient, except for the all-ones
significand (a tie), which the kernel corrects when the device check (hardware reciprocal + step
against the IEEE quotient, run when a context is created) does not find the step exact.  Scaling
y by 2^k scales every quantity exactly, so the result holds for every normal y.
"""
import numpy as np


def _fma32(a, x, c):
    """RN32(a x + c) for float32 operands, exactly: the product is exact in float64, the sum's
    error is recovered with TwoSum, and a float64 sum that falls on a float32 midpoint is resolved
    by the sign of that error."""
    a = a.astype(np.float64)
    x = x.astype(np.float64)
    c = c.astype(np.float64)
    p = a * x
    hi = p + c
    bb = hi - p
    lo = (p - (hi - bb)) + (c - bb)
    r = hi.astype(np.float32)
    r64 = r.astype(np.float64)
    toward = np.where(hi > r64, np.float32(np.inf), np.float32(-np.inf))
    nxt = np.nextafter(r, toward).astype(np.float64)
    on_mid = (hi == (r64 + nxt) / 2) & (r64 != hi)
    to_next = np.where(nxt > r64, lo > 0, lo < 0)
    return np.where(on_mid & to_next, nxt.astype(np.float32), r)


def test_fma32_emulation_rounds_like_float32():
    # against exact rational arithmetic on a seeded sample
    from fractions import Fraction
    rng = np.random.default_rng(5)
    a = rng.uniform(0.5, 2, 400).astype(np.float32)
    x = rng.uniform(-2, 2, 400).astype(np.float32)
    c = (rng.uniform(-1, 1, 400) * 2.0 ** rng.integers(-30, 2, 400)).astype(np.float32)
    got = _fma32(a, x, c)
    for i in range(400):
        exact = Fraction(float(a[i])) * Fraction(float(x[i])) + Fraction(float(c[i]))
        lo = np.float32(float(exact))
        # the float32 neighbours of the exact value; pick the nearest, ties to even
        cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
        best = min(cands, key=lambda f: (abs(Fraction(float(f)) - exact), int(np.float32(f).view(np.uint32)) & 1))
        assert got[i] == best, (a[i], x[i], c[i])


def test_newton_step_is_correctly_rounded():
    m = np.arange(1 << 23, dtype=np.uint32)
    y = (np.uint32(127 << 23) | m).view(np.float32)
    q = np.float32(1.0) / y  # IEEE quotient
    below = np.where(q.astype(np.float64) <= 1.0 / y.astype(np.float64), q, np.nextafter(q, np.float32(0)))
    above = np.nextafter(below, np.float32(2))
    all_ones = m == 0x7FFFFF
    for r in (below, above):
        e = _fma32(-y, r, np.ones_like(r))
        refined = _fma32(r, e, r)
        bad = (refined != q) & ~all_ones
        assert not bad.any(), y[bad][:4]
    # the excluded significand does tie the wrong way from below; setting the lowest significand
    # bit (what the kernel does there when the device check fails) gives the quotient from both
    qa = q[all_ones]
    assert (qa.view(np.uint32) & 0x7FFFFF)[0] == 1
    for r in (below[all_ones], above[all_ones]):
        e = _fma32(-y[all_ones], r, np.ones(1, np.float32))
        refined = _fma32(r, e, r)
        assert (refined.view(np.uint32) | np.uint32(1)).view(np.float32)[0] == qa[0]
    e = _fma32(-y[all_ones], below[all_ones], np.ones(1, np.float32))
    assert _fma32(below[all_ones], e, below[all_ones])[0] != qa[0]
