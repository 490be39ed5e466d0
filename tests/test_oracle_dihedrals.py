"""Pins the oracle's dihedral and pseudo-site geometry by hand-derived closed forms.

Each case places two nucleotides in a geometry simple enough to write every argument
of the term by hand (a twist about the stacking axis, a tilt, a lateral offset),
derived from the term's definition in the reference:

  * oxDNA stacking (src/OxDNABondedKernel.h:81-126): p = B (the 5' neighbour), q = A;
    cos(phi1) = a2 . r_bb / |r_bb|, cos(phi2) = b2 . r_bb / |r_bb| with r_bb from B's
    back reference site to A's, theta4 from a3 . b3, theta5 = acos(-a3 . r_st),
    theta6 = acos(-b3 . r_st), r_st from B's stack site to A's; epsilon st_eps[A][B];
  * oxRNA stacking (:250-298): STACK_3 on B, STACK_5 on A, thetaB1 / thetaB2 from
    BBVECTOR_3 (B) / BBVECTOR_5 (A) against the real BACK-BACK direction, phi1 / phi2
    from that same vector, no theta4;
  * coaxial stacking (src/EvaluatorPairOxDNACoaxStacking.h:150-215): cos(phi3) =
    r_st . (r_br x a1) / (|r_st| |r_br|), and for oxRNA cos(phi4) with b1 in place of
    a1, r_br between the back reference sites.

The expected energy is then the product of the term's radial factor and its f4 / f5
factors at those arguments, each evaluated by the oracle's primitives (pinned on their
own by closed forms in test_oracle_primitives.py). The dihedral cosines are chosen
negative where the case pins them: for cos(phi) >= 0 the factor f5 is identically 1,
which would pin nothing. Plausible mistakes each change one of the arguments and so the
energy: a1 used for a2 in phi1 / phi2, the stack site for the back reference in r_bb,
the sign of a dihedral, STACK_3 / STACK_5 or BBVECTOR_3 / BBVECTOR_5 exchanged, the
particles' roles in thetaB1 / thetaB2, b1 replaced by a1 in phi4.
"""

import math

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import _abi
from oxdna.model import dna1, dna2, rna2

EX, EY, EZ = np.eye(3)


def quat_of(a1, a3):
    return H.axes_quat(np.asarray(a1, float), np.asarray(a3, float))


def f1(r, P, eps):
    params = [P.st_a, P.st_r0, P.st_rlow, P.st_rhigh, P.st_rclow, P.st_rchigh, P.st_blow, P.st_bhigh,
              eps, eps * P.st_shift_factor]
    return O.prim(0, r, params)[0]


def f2(r, p):
    return O.prim(1, r, [p.f2_k, p.f2_r0, p.f2_rc, p.f2_rlow, p.f2_rhigh, p.f2_rclow, p.f2_rchigh, p.f2_blow,
                         p.f2_bhigh])[0]


def f4(theta, s):
    return O.prim(2, theta, list(s))[0]


def f5(x, s):
    return O.prim(3, x, list(s))[0]


def stacking(model, **kw):
    return model.bonded_force(**kw)._pack(["backbone"])[0]


def coax(model):
    force = model.coaxial_stacking(model.make_neighbor_list())
    table, rcut = force._pack(["A", "G", "C", "T"])
    return table[0], rcut[0]


def bond_energy(P, rB, qA, qB, tA=0, tB=0):
    """Stacking energy of the bond (A at the origin, B = its 5' neighbour at rB)."""
    return O.bond(P, rB, qA, qB, tA, tB)["e_stack"]


def eps_of(P, tA=0, tB=0):
    return P.st_eps[4 * tA + tB]


# ------------------------------------------------------------------ oxDNA stacking
@pytest.mark.parametrize("model", [dna1, dna2])
def test_dna_stacking_twist(model):
    """B above A by h along the common a3 = z, turned by alpha about it: a1 = (cos, sin, 0),
    a2 = (-sin, cos, 0). With the back reference at -0.4 a1 and the stack site at 0.34 a1:
      r_bb = (-0.4 (1 - cos a), 0.4 sin a, -h),   cos(phi1) = cos(phi2) = 0.4 sin a / |r_bb|,
      r_st = (0.34 (1 - cos a), -0.34 sin a, -h), theta4 = 0, theta5 = theta6 = acos(h / |r_st|)."""
    P = stacking(model)
    assert tuple(P.stack_back_ref) == (-0.4, 0.0, 0.0) and tuple(P.stack_site) == (0.34, 0.0, 0.0)
    h = 0.39
    for alpha in (-0.5, -0.3):
        ca, sa = math.cos(alpha), math.sin(alpha)
        r_bb = math.sqrt(0.32 * (1 - ca) + h * h)
        r_st = math.sqrt(2 * 0.34 ** 2 * (1 - ca) + h * h)
        cphi = 0.4 * sa / r_bb
        assert -0.65 < cphi < 0.0  # the quadratic branch of f5
        th56 = math.acos(h / r_st)
        expect = (f1(r_st, P, eps_of(P)) * f4(0.0, P.t4) * f4(th56, P.t5) ** 2 * f5(cphi, P.p1) ** 2)
        assert abs(expect) > 1e-3
        E = bond_energy(P, h * EZ, quat_of(EX, EZ), quat_of((ca, sa, 0), EZ))
        assert E == pytest.approx(expect, rel=1e-12, abs=1e-14), alpha


@pytest.mark.parametrize("model", [dna1, dna2])
def test_dna_stacking_tilt(model):
    """B above A by h, its frame tilted by tau about x: a2 = (0, cos t, sin t), a3 = (0, -sin t,
    cos t). The sites stay on the axis: r_bb = r_st = (0, 0, -h), so cos(phi1) = -sin t,
    cos(phi2) = 0, theta4 = theta5 = tau, theta6 = 0."""
    P = stacking(model)
    h = 0.4
    for tau in (0.3, 0.45):
        ct, st = math.cos(tau), math.sin(tau)
        expect = (f1(h, P, eps_of(P)) * f4(tau, P.t4) * f4(tau, P.t5) * f4(0.0, P.t5)
                  * f5(-st, P.p1) * f5(0.0, P.p1))
        assert abs(expect) > 1e-3 and f5(-st, P.p1) < 0.99
        E = bond_energy(P, h * EZ, quat_of(EX, EZ), quat_of(EX, (0, -st, ct)))
        assert E == pytest.approx(expect, rel=1e-12, abs=1e-14), tau


def test_dna_stacking_twist_and_tilt():
    """Frame of B = Rz(alpha) Rx(tau): a1 = (cos a, sin a, 0), a2 = (-sin a cos t, cos a cos t,
    sin t), a3 = (sin a sin t, -cos a sin t, cos t). Then cos(phi1) = (0.4 sin a cos t -
    h sin t) / |r_bb| but cos(phi2) = 0.4 sin a / |r_bb|, theta4 = tau, theta6 = acos(h /
    |r_st|), theta5 = acos((h cos t - 0.34 sin a sin t) / |r_st|)."""
    P = stacking(dna2)
    h, alpha, tau = 0.4, -0.35, 0.2
    ca, sa, ct, st = math.cos(alpha), math.sin(alpha), math.cos(tau), math.sin(tau)
    a1, a3 = (ca, sa, 0.0), (sa * st, -ca * st, ct)
    r_bb = math.sqrt(0.32 * (1 - ca) + h * h)
    r_st = math.sqrt(2 * 0.34 ** 2 * (1 - ca) + h * h)
    cphi1 = (0.4 * sa * ct - h * st) / r_bb
    cphi2 = 0.4 * sa / r_bb
    th5 = math.acos((h * ct - 0.34 * sa * st) / r_st)
    th6 = math.acos(h / r_st)
    expect = (f1(r_st, P, eps_of(P)) * f4(tau, P.t4) * f4(th5, P.t5) * f4(th6, P.t5)
              * f5(cphi1, P.p1) * f5(cphi2, P.p1))
    assert abs(expect) > 1e-3
    E = bond_energy(P, h * EZ, quat_of(EX, EZ), quat_of(a1, a3))
    assert E == pytest.approx(expect, rel=1e-12)


def test_dna_stacking_direction():
    """Stacking is directional: with B (the 5' neighbour) below A instead of above it,
    theta5 = acos(-a3 . r_st) = pi and f4 vanishes."""
    P = stacking(dna2)
    q = quat_of(EX, EZ)
    assert bond_energy(P, 0.4 * EZ, q, q) < -0.1
    assert bond_energy(P, -0.4 * EZ, q, q) == 0.0


def test_dna_stacking_epsilon_index_order():
    """The sequence-dependent epsilon is st_eps[type of A][type of B]: the energy at one
    geometry scales with it (radial factor and shift are both proportional), and the
    oxDNA2 table is not symmetric."""
    P = stacking(dna2, average=False)
    q = quat_of(EX, EZ)
    e_ref = bond_energy(P, 0.4 * EZ, q, q, 0, 0)
    asym = 0
    for tA in range(4):
        for tB in range(4):
            E = bond_energy(P, 0.4 * EZ, q, q, tA, tB)
            assert E == pytest.approx(e_ref * eps_of(P, tA, tB) / eps_of(P, 0, 0), rel=1e-12)
            asym += eps_of(P, tA, tB) != eps_of(P, tB, tA)
    assert asym > 0


# ------------------------------------------------------------------ oxRNA stacking
def _rna_expect(P, h, tau=0.0, alpha=0.0):
    """Closed forms for the two oxRNA geometries below (one of tau / alpha nonzero)."""
    s3, s5 = np.array(P.stack_3_site), np.array(P.stack_5_site)
    b3, b5 = np.array(P.bbvector_3), np.array(P.bbvector_5)
    assert s3[2] == 0.0 and s5[2] == 0.0
    assert tuple(P.back_site) == (-0.4, 0.0, 0.2)
    if alpha != 0.0:
        # B turned by alpha about z (a3 = b3 = z): the back sites' z parts cancel
        ca, sa = math.cos(alpha), math.sin(alpha)
        Rs3 = np.array([s3[0] * ca - s3[1] * sa, s3[0] * sa + s3[1] * ca])
        r_st = math.sqrt(h * h + float(np.sum((s5[:2] - Rs3) ** 2)))
        r_bb = math.sqrt(0.32 * (1 - ca) + h * h)
        cphi1 = cphi2 = 0.4 * sa / r_bb
        ctB1 = (h * b3[2] - 0.4 * (b3[0] * (1 - ca) + b3[1] * sa)) / r_bb
        ctB2 = (0.4 * (1 - ca) * b5[0] - 0.4 * sa * b5[1] + h * b5[2]) / r_bb
        th5 = th6 = math.acos(h / r_st)
    else:
        # B tilted by tau about x: a2 = (0, cos, sin), a3 = (0, -sin, cos)
        ct, st = math.cos(tau), math.sin(tau)
        r_st = math.sqrt((s5[0] - s3[0]) ** 2 + (s5[1] - s3[1] * ct) ** 2 + (h + s3[1] * st) ** 2)
        bz = 0.2 - h - 0.2 * ct                       # r_bb = (0, 0.2 sin t, bz)
        r_bb = math.sqrt((0.2 * st) ** 2 + bz * bz)
        cphi1 = st * (0.2 - h) / r_bb
        cphi2 = 0.2 * st / r_bb
        Rb3y, Rb3z = b3[1] * ct - b3[2] * st, b3[1] * st + b3[2] * ct
        ctB1 = -(0.2 * st * Rb3y + bz * Rb3z) / r_bb
        ctB2 = -(0.2 * st * b5[1] + bz * b5[2]) / r_bb
        th5 = math.acos((s5[1] * st + h * ct) / r_st)
        th6 = math.acos((h + s3[1] * st) / r_st)
    return (f1(r_st, P, eps_of(P)) * f4(th5, P.t5) * f4(th6, P.t5) * f4(math.acos(ctB1), P.tB1)
            * f4(math.acos(ctB2), P.tB2) * f5(cphi1, P.p1) * f5(cphi2, P.p1)), cphi1, cphi2


def test_rna_stacking_twist():
    """B above A by h, turned by alpha about z: thetaB1 / thetaB2 from BBVECTOR_3 rotated with
    B and BBVECTOR_5 with A against r_bb = (-0.4 (1 - cos a), 0.4 sin a, -h); r_st from
    STACK_3 (rotated with B) to STACK_5."""
    P = stacking(rna2)
    assert P.rna_stacking == 1
    h = 0.6
    for alpha in (0.3, 0.45):
        expect, _, _ = _rna_expect(P, h, alpha=alpha)
        assert abs(expect) > 1e-3
        E = bond_energy(P, h * EZ, quat_of(EX, EZ), quat_of((math.cos(alpha), math.sin(alpha), 0), EZ))
        assert E == pytest.approx(expect, rel=1e-12), alpha


def test_rna_stacking_tilt():
    """B above A by h, tilted by tau about x: cos(phi1) = sin t (0.2 - h) / |r_bb| < 0 pins phi1
    (phi2 positive); thetaB1 from BBVECTOR_3 tilted with B. The model's BBVECTORs put both
    thetaB near their cut-off in this geometry, so the case carries unit vectors of its own
    nearer the back-back direction (what is pinned is which particle turns which vector
    and what it is dotted with)."""
    P = stacking(rna2)
    for name, v in (("bbvector_3", (0.1, 0.2, 0.97)), ("bbvector_5", (-0.1, -0.15, 0.98))):
        v = np.asarray(v) / np.linalg.norm(v)
        getattr(P, name)[:] = v
    h = 0.4
    for tau in (0.25, 0.35):
        expect, cphi1, cphi2 = _rna_expect(P, h, tau=tau)
        assert abs(expect) > 1e-3 and -0.65 < cphi1 < 0.0
        ct, st = math.cos(tau), math.sin(tau)
        E = bond_energy(P, h * EZ, quat_of(EX, EZ), quat_of(EX, (0, -st, ct)))
        assert E == pytest.approx(expect, rel=1e-12), tau


# ------------------------------------------------------------------ coaxial stacking
def _coax_expect(p, h, d, tau, back_z):
    """i at the origin (identity frame); j at (0, d, h) with frame Ry(tau): b1 = (cos t, 0,
    -sin t), b3 = (sin t, 0, cos t). Stack site 0.34 a1; back reference -0.4 a1 + back_z a3.
      r_st = (0.34 (cos t - 1), d, h - 0.34 sin t)
      r_br = (0.4 (1 - cos t) + bz sin t, d, h + 0.4 sin t + bz (cos t - 1))
      |r_st| |r_br| cos(phi3) = d (0.74 sin t + bz (cos t - 1))
      |r_st| |r_br| cos(phi4) = d (0.74 sin t + bz (1 - cos t))
    theta1 = pi - tau, theta4 = tau, cos(theta5) = (h - 0.34 sin t) / |r_st|,
    cos(theta6) = (0.34 sin t - h cos t) / |r_st|."""
    ct, st = math.cos(tau), math.sin(tau)
    bz = back_z
    rs = np.array([0.34 * (ct - 1), d, h - 0.34 * st])
    rb = np.array([0.4 * (1 - ct) + bz * st, d, h + 0.4 * st + bz * (ct - 1)])
    ns, nb = math.sqrt(float(rs @ rs)), math.sqrt(float(rb @ rb))
    cphi3 = d * (0.74 * st + bz * (ct - 1)) / (ns * nb)
    cphi4 = d * (0.74 * st + bz * (1 - ct)) / (ns * nb)
    th1 = math.pi - tau
    g1 = f4(th1, p.t1)
    if p.t1_mode == 0:
        g1 += f4(2 * math.pi - th1, p.t1)
    elif th1 > p.t1_sb:
        g1 += p.t1_sa * (th1 - p.t1_sb) ** 2
    th5 = math.acos((h - 0.34 * st) / ns)
    th6 = math.acos((0.34 * st - h * ct) / ns)
    g5 = f4(th5, p.t56) + f4(math.pi - th5, p.t56)
    g6 = f4(th6, p.t56) + f4(math.pi - th6, p.t56)
    if p.rna_coax:
        phi = f5(cphi3, p.phi3) * f5(cphi4, p.phi4)
    elif p.phi3_enabled:
        phi = f5(cphi3, p.phi3) ** 2
    else:
        phi = 1.0
    return f2(ns, p) * g1 * f4(tau, p.t4) * g5 * g6 * phi, cphi3, cphi4


@pytest.mark.parametrize("model,back_z,h,d", [(dna1, 0.0, 0.4, -0.15), (dna2, 0.0, 0.4, -0.15),
                                              (rna2, 0.2, 0.55, -0.25)])
def test_coax_tilt_and_offset(model, back_z, h, d):
    p, rc = coax(model)
    assert tuple(p.stack_site) == (0.34, 0.0, 0.0)
    assert tuple(p.stack_back_ref) == (-0.4, 0.0, back_z)
    for tau in (0.2, 0.3):
        expect, cphi3, cphi4 = _coax_expect(p, h, d, tau, back_z)
        assert abs(expect) > 1e-3
        if model is not dna2:
            assert -0.65 < cphi3 < 0.0 and -0.65 < cphi4 < 0.0
        if model is rna2:
            assert abs(cphi3 - cphi4) > 0.005  # phi4 (b1) is not phi3 (a1) here
        ct, st = math.cos(tau), math.sin(tau)
        rj = np.array([0.0, d, h])
        _, _, _, _, E, _ = O.pair(_abi.TERM_COAX, -rj, quat_of(EX, EZ), quat_of((ct, 0, -st), (st, 0, ct)), rc, p)
        assert E == pytest.approx(expect, rel=1e-12), (model.__name__, tau)


def test_coax_dna1_twist():
    """j above i by h, turned by alpha about z: |r_st| |r_br| cos(phi3) = 0.74 h sin a,
    theta1 = pi - |alpha|, theta4 = 0, theta6 = pi - theta5."""
    p, rc = coax(dna1)
    h = 0.4
    for alpha in (-0.3, -0.2):
        ca, sa = math.cos(alpha), math.sin(alpha)
        ns = math.sqrt(2 * 0.34 ** 2 * (1 - ca) + h * h)
        nb = math.sqrt(2 * 0.4 ** 2 * (1 - ca) + h * h)
        cphi3 = 0.74 * h * sa / (ns * nb)
        assert -0.65 < cphi3 < 0.0
        th1 = math.pi - abs(alpha)
        th5 = math.acos(h / ns)
        g5 = f4(th5, p.t56) + f4(math.pi - th5, p.t56)
        expect = (f2(ns, p) * (f4(th1, p.t1) + f4(2 * math.pi - th1, p.t1)) * f4(0.0, p.t4) * g5 * g5
                  * f5(cphi3, p.phi3) ** 2)
        assert abs(expect) > 1e-3
        _, _, _, _, E, _ = O.pair(_abi.TERM_COAX, -h * EZ, quat_of(EX, EZ), quat_of((ca, sa, 0), EZ), rc, p)
        assert E == pytest.approx(expect, rel=1e-12), alpha
