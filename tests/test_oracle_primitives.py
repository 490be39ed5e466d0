"""Pins the oracle's building blocks (reference src/OxDNANumerics.h) by mathematics:
closed forms at sample points and value / slope continuity at every smoothing knot
of every parameter set the models use. The knots' constants (b, r_c, ...) are
derived by oxDNA to make each function C1, so a wrong constant in the restated model
tables or a wrong branch formula breaks a continuity check."""

import math

import numpy as np
import pytest

import oracle as O
from oxdna.model import dna1, dna2, rna2

H_STEP = 1e-9
# oxDNA publishes the smoothing constants to ~6 significant digits, so value and slope
# match across a knot to ~1e-5 relative; a wrong constant or branch formula gives
# jumps of 1e-2 or more.
KNOT_REL = 2e-5


def f1_params(shape, eps, shift_factor):
    a, r0, rlow, rhigh, rclow, rchigh, blow, bhigh = shape
    return [a, r0, rlow, rhigh, rclow, rchigh, blow, bhigh, eps, eps * shift_factor]


F1_SETS = {
    "hb_dna": (dna1._hb_f1(), 1.077, dna1.hb_shift_factor()),
    "hb_rna": (rna2._hb_f1(), 0.870439, rna2._hb_shift()),
    "stack_dna": ((dna1.STCK_A, dna1.STCK_R0, dna1.STCK_RLOW, dna1.STCK_RHIGH, dna1.STCK_RCLOW, dna1.STCK_RCHIGH,
                   dna1.STCK_BLOW, dna1.STCK_BHIGH), 1.6, dna1._stack_shift_factor()),
    "stack_rna": ((rna2.STCK_A, rna2.STCK_R0, rna2.STCK_RLOW, rna2.STCK_RHIGH, rna2.STCK_RCLOW, rna2.STCK_RCHIGH,
                   rna2.STCK_BLOW, rna2.STCK_BHIGH), 1.6,
                  dna1._morse_shift(rna2.STCK_A, rna2.STCK_RC, rna2.STCK_R0)),
}
F2_SETS = {"crst_dna": dna1.CRST_F2, "cxst_dna1": dna1.CXST_F2, "cxst_dna2": (dna2.CXST_K,) + dna1.CXST_F2[1:],
           "crst_rna": rna2.CRST_F2, "cxst_rna": rna2.CXST_F2}
F4_SETS = {
    "hb_t1": dna1.HB_F4_THETA1, "hb_t4": dna1.HB_F4_THETA4, "hb_t78": dna1.HB_F4_THETA78,
    "cs_t1": dna1.CRST_F4_THETA1, "cs_t23": dna1.CRST_F4_THETA23, "cs_t4": dna1.CRST_F4_THETA4,
    "cs_t78": dna1.CRST_F4_THETA78, "st_t4": dna1.STCK_F4_THETA4, "st_t5": dna1.STCK_F4_THETA5,
    "cx_t1": dna1.CXST_F4_THETA1, "cx2_t1": dna2.CXST_F4_THETA1, "rna_cs_t1": rna2.CRST_F4_THETA1,
    "rna_cx_t1": rna2.CXST_F4_THETA1, "rna_cx_t4": rna2.CXST_F4_THETA4, "rna_cx_t56": rna2.CXST_F4_THETA56,
    "rna_st_tB1": rna2.STCK_F4_THETAB1,
}
F5_SETS = {"st_phi1": dna1.STCK_F5_PHI1, "cx_phi3": dna1.CXST_F5_PHI3, "rna_phi": rna2.STCK_F5_PHI1}


def cont(which, x, p, slope_scale=None):
    """Value and slope (or, for f4, slope = dsin * sin) agree on both sides of a knot."""
    lo, hi = O.prim(which, x - H_STEP, p), O.prim(which, x + H_STEP, p)
    s = math.sin(x) if which == 2 else 1.0
    assert abs(lo[0] - hi[0]) <= KNOT_REL * max(1.0, abs(lo[0])), (which, x, lo, hi)
    assert abs(lo[1] * s - hi[1] * s) <= KNOT_REL * max(1.0, abs(lo[1] * s)), (which, x, lo, hi)
    return lo, hi


@pytest.mark.parametrize("name", sorted(F1_SETS))
def test_f1_morse_closed_form_and_c1(name):
    shape, eps, sf = F1_SETS[name]
    p = f1_params(shape, eps, sf)
    a, r0, rlow, rhigh, rclow, rchigh = shape[:6]
    # closed form inside the Morse branch, and the minimum at r0
    for r in np.linspace(rlow + 1e-3, rhigh - 1e-3, 7):
        v, d = O.prim(0, r, p)
        assert v == pytest.approx(eps * (1 - math.exp(-(r - r0) * a)) ** 2 - eps * sf, rel=1e-12, abs=1e-14)
        assert d == pytest.approx(2 * eps * a * math.exp(-(r - r0) * a) * (1 - math.exp(-(r - r0) * a)),
                                  rel=1e-12, abs=1e-14)
    assert O.prim(0, r0, p) == pytest.approx((-eps * sf, 0.0), abs=1e-14)
    # value and slope continuous at the four knots; zero outside [rclow, rchigh]
    for knot in (rclow, rlow, rhigh, rchigh):
        cont(0, knot, p)
    assert O.prim(0, rchigh + 1e-3, p) == (0.0, 0.0)
    assert O.prim(0, rclow - 1e-3, p) == (0.0, 0.0)
    # the derivative is the slope of the value (central difference)
    for r in np.linspace(rclow + 1e-3, rchigh - 1e-3, 25):
        h = 1e-6
        fd = (O.prim(0, r + h, p)[0] - O.prim(0, r - h, p)[0]) / (2 * h)
        assert O.prim(0, r, p)[1] == pytest.approx(fd, rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("name", sorted(F2_SETS))
def test_f2_harmonic_closed_form_and_c1(name):
    p = list(F2_SETS[name])
    k, r0, rc, rlow, rhigh, rclow, rchigh = p[:7]
    for r in np.linspace(rlow + 1e-3, rhigh - 1e-3, 5):
        v, d = O.prim(1, r, p)
        assert v == pytest.approx(0.5 * k * ((r - r0) ** 2 - (rc - r0) ** 2), rel=1e-12, abs=1e-14)
        assert d == pytest.approx(k * (r - r0), rel=1e-12, abs=1e-14)
    for knot in (rclow, rlow, rhigh, rchigh):
        cont(1, knot, p)
    for r in np.linspace(rclow + 1e-3, rchigh - 1e-3, 25):
        h = 1e-6
        fd = (O.prim(1, r + h, p)[0] - O.prim(1, r - h, p)[0]) / (2 * h)
        assert O.prim(1, r, p)[1] == pytest.approx(fd, rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("name", sorted(F4_SETS))
def test_f4_angular_closed_form_and_c1(name):
    p = list(F4_SETS[name])
    a, b, t0, ts, tc = p
    assert O.prim(2, t0, p)[0] == pytest.approx(1.0)
    for dt in (0.3 * ts, -0.5 * ts):
        th = t0 + dt
        if 0 < th < math.pi:
            v, dsin = O.prim(2, th, p)
            assert v == pytest.approx(1 - a * dt * dt, rel=1e-12)
            assert dsin * math.sin(th) == pytest.approx(-2 * a * dt, rel=1e-9)
    for knot in (t0 + ts, t0 + tc, t0 - ts, t0 - tc):
        if not (1e-3 < knot < math.pi - 1e-3):
            continue
        cont(2, knot, p)
    # dsin * sin equals the slope of the value
    for th in np.linspace(0.05, math.pi - 0.05, 40):
        h = 1e-6
        fd = (O.prim(2, th + h, p)[0] - O.prim(2, th - h, p)[0]) / (2 * h)
        assert O.prim(2, th, p)[1] * math.sin(th) == pytest.approx(fd, rel=1e-5, abs=1e-5)


def test_f4dsin_small_sine_fallback_is_the_limit():
    """Where sin(theta)^2 <= 1e-8 the inner branch returns -m 2a instead of
    -m 2a (theta - t0) / sin(theta) (reference src/OxDNANumerics.h:236-241). The inner
    branch reaches sin(theta) -> 0 only for t0 at 0 or pi, and there the ratio tends to
    -2a (t0 = 0, theta -> 0+) or +2a (t0 = pi, theta -> pi-): the fallback must equal
    that limit, and the values just outside the guard must tend to it."""
    sets = {n: list(p) for n, p in F4_SETS.items() if min(abs(p[2]), abs(p[2] - math.pi)) < 1e-9}
    assert {"hb_t1", "hb_t4"} <= set(sets)
    for name, p in sets.items():
        a, b, t0, ts, tc = p
        side = 1.0 if abs(t0) < 1e-9 else -1.0  # the side of t0 inside [0, pi]
        limit = -2 * a if side > 0 else 2 * a
        for eps in (1e-6, 5e-5, 9e-5):  # sin^2 <= 1e-8: the fallback
            th = t0 + side * eps
            assert math.sin(th) ** 2 <= 1e-8
            assert O.prim(2, th, p)[1] == pytest.approx(limit, rel=1e-12), name
        for eps in (1.01e-4, 3e-4, 1e-3):  # outside the guard: -m 2a eps / sin(eps)
            th = t0 + side * eps
            assert math.sin(th) ** 2 > 1e-8
            assert O.prim(2, th, p)[1] == pytest.approx(limit, rel=2e-7), name


@pytest.mark.parametrize("name", sorted(F5_SETS))
def test_f5_dihedral_closed_form_and_c1(name):
    p = list(F5_SETS[name])
    a, b, xc, xs = p
    assert O.prim(3, 0.3, p) == (1.0, 0.0)
    assert O.prim(3, xc - 1e-3, p) == (0.0, 0.0)
    c = 0.5 * xs
    assert O.prim(3, c, p)[0] == pytest.approx(1 - a * c * c)
    for knot in (xc, xs, 0.0):
        cont(3, knot, p)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_excluded_volume_lj_closed_form_and_c1(k):
    sigma, rstar, b, rc = dna1.EXCL[k]
    p = [dna1.EXCL_EPS, sigma, rstar, b, rc]
    eps = dna1.EXCL_EPS
    for r in (0.8 * sigma, 0.95 * rstar, 0.99 * rstar):  # LJ branch: r < r*
        v, f = O.prim(4, r, p)
        lj = 4 * eps * ((sigma / r) ** 12 - (sigma / r) ** 6)
        assert v == pytest.approx(lj, rel=1e-12, abs=1e-12)
        dlj = 4 * eps * (-12 * sigma ** 12 / r ** 13 + 6 * sigma ** 6 / r ** 7)
        assert f == pytest.approx(-dlj, rel=1e-10)  # force on the second site along +x
    for knot in (rstar, rc):
        cont(4, knot, p)
    assert O.prim(4, rc + 1e-6, p) == (0.0, 0.0)


def test_fene_closed_form():
    eps, r0, delta = dna1.FENE_EPS, dna1.FENE_R0, dna1.FENE_DELTA
    for dr in (-0.2, -0.05, 0.0, 0.1, 0.24):
        r = r0 + dr
        v, f = O.prim(5, r, [eps, r0, delta])
        assert v == pytest.approx(-0.5 * eps * math.log(1 - dr * dr / (delta * delta)), rel=1e-12, abs=1e-15)
        # force on the second site = -dE/dr
        assert f == pytest.approx(-eps * dr / (delta * delta - dr * dr), rel=1e-12, abs=1e-15)
