"""The cancellation-free short-range law of the PRECISE PP mode, on the CPU.

For 1 < ep <= 2 (ep = 2r/a) the S2 force law of basic_functions.f08:95-109
satisfies 35 a^2 ep^2 (F_ra - 1/r^2) = (ep-2)^5 (7ep^3 + 22ep^2 + 10ep + 4),
so 1/r^2 - F_ra = (2-ep)^5 (7ep^3+22ep^2+10ep+4) / (140 r^2) without the
cancellation of the expanded form.  Checked here as an integer polynomial
identity and by a float32 emulation of the kernel's evaluation order
against FP64.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import polynomial as P

f = np.float32


def test_factorisation_is_an_integer_identity():
    # 35 a^2 ep^2 F_ra for 1 < ep <= 2 as coefficients of ep^0..ep^8; with
    # 1/r^2 = 4/(a^2 ep^2), 35 a^2 ep^2 / r^2 = 140
    expanded = [12, 0, -224, 896, -840, 224, 70, -48, 7]
    minus_140 = list(expanded)
    minus_140[0] -= 140                     # 35 a^2 ep^2 (F_ra - 1/r^2)
    factored = P.polymul(P.polypow([-2, 1], 5), [4, 10, 22, 7])
    assert np.array_equal(np.round(factored).astype(int), minus_140)
    # continuity with the inner branch at ep = 1: both give 35 a^2 (F_ra - 1/r^2) = -43
    inner = [0, 224, 0, -224, 70, 48, -21]       # 35 a^2 F_ra for ep <= 1
    assert P.polyval(1, inner) - 140 == -43
    assert P.polyval(1, factored) == -43


def _fra64(r, a):
    ep = 2 * r / a
    hi = (12 / ep ** 2 - 224 + 896 * ep - 840 * ep ** 2 + 224 * ep ** 3 + 70 * ep ** 4
          - 48 * ep ** 5 + 7 * ep ** 6) / (35 * a * a)
    lo = (224 * ep - 224 * ep ** 3 + 70 * ep ** 4 + 48 * ep ** 5 - 21 * ep ** 6) / (35 * a * a)
    return np.where(ep > 2, 1 / r ** 2, np.where(ep > 1, hi, lo))


def _short_f32(r, a):
    """The kernel's evaluation of 1/r^2 - F_ra(r, a), op for op in float32."""
    ia2 = f(2.0 / np.float64(a))
    ep = r * ia2
    ir = f(1) / r
    ir2 = ir * ir
    t = f(2) - ep
    t2 = t * t
    q = ep * (ep * (ep * f(7) + f(22)) + f(10)) + f(4)
    hi = t2 * t2 * t * q * ir2 * f(1 / 140)
    c = f(1 / 35 / (np.float64(a) ** 2))
    e2 = ep * ep
    lo = ir2 - c * ep * (e2 * (ep * (ep * (ep * f(-21) + f(48)) + f(70)) + f(-224)) + f(224))
    return np.where(ep > 1, hi, lo)


def _fra_f32_expanded(r, a):
    """F_ra term by term in float32, the expanded form's rounding."""
    ep = f(2) * r / a
    e2 = ep * ep
    e3, e4 = e2 * ep, e2 * e2
    e5, e6 = e4 * ep, e3 * e3
    hi = (f(12) / e2 - f(224) + f(896) * ep - f(840) * e2 + f(224) * e3 + f(70) * e4
          - f(48) * e5 + f(7) * e6) / (f(35) * a * a)
    lo = (f(224) * ep - f(224) * e3 + f(70) * e4 + f(48) * e5 - f(21) * e6) / (f(35) * a * a)
    return np.where(ep > 1, hi, lo)


def test_factored_law_error_is_far_below_the_expanded_form():
    with np.errstate(all="ignore"):
        for a in (f(3.5 / 2), f(3.5 / 4), f(0.1)):
            rng = np.random.default_rng(1)
            r = (rng.uniform(0, 1, 200000) ** (1 / 3) * np.float64(a)).astype(f)   # uniform in volume
            r = r[r > 0]
            r64 = r.astype(np.float64)
            exact = 1 / r64 ** 2 - _fra64(r64, np.float64(a))
            err_short = np.abs(_short_f32(r, a).astype(np.float64) - exact) * r64 ** 2
            err_exp = np.abs(_fra_f32_expanded(r, a).astype(np.float64)
                             - _fra64(r64, np.float64(a))) * r64 ** 2
            rms_s = np.sqrt((err_short ** 2).mean())
            rms_e = np.sqrt((err_exp ** 2).mean())
            # errors in units of 1/r^2, the scale of each pair's terms
            assert rms_s < 1e-7 and err_short.max() < 1e-6
            assert rms_s < rms_e / 20
