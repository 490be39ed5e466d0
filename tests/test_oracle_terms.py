"""Pins the oracle's per-term physics.

  * closed-form special geometries: every angle of a term at its optimum reduces the
    energy to its radial factor (HB, stacking, coaxial), a single active site pair
    reduces excluded volume to textbook LJ, Debye-Hueckel to a screened Coulomb term;
    a rotation that moves exactly one angle scans exactly one f4 factor;
  * the symmetrised cross-stacking factors (f4(t) + f4(pi - t)) on t4 / t7 / t8 with
    f4 == 1 give 2^3 (or 2^2 with t4 dropped) times the radial factor;
  * forces and torques equal central finite differences of the energy (the
    reference's fd_check method, pytest/_helpers.py:137-181, with an FP64 step);
  * Newton's third law and total angular momentum conservation per pair / system;
  * fused HB + cross-stacking == HB and cross-stacking evaluated apart.
"""

import math

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import _abi, forces as FF, io, systems
from oxdna.model import dna1, dna2, rna2

EX, EY, EZ = np.eye(3)


def first_struct(force, pair=("A", "T")):
    table, rcut = force._pack(["A", "G", "C", "T"])
    i, j = ["A", "G", "C", "T"].index(pair[0]), ["A", "G", "C", "T"].index(pair[1])
    return table[i * 4 + j], rcut[i * 4 + j]


def ident():
    return np.array([1.0, 0.0, 0.0, 0.0])


# ------------------------------------------------------------------ closed forms
def ideal_hb_pair(r):
    """i at the origin (a1 = x, a3 = z); j placed so that every HB angle sits at its
    optimum: b1 = -a1, b3 = -a3, base-base vector along a1 with length r."""
    qi = H.axes_quat(EX, EZ)
    qj = H.axes_quat(-EX, -EZ)
    pos_j = (0.4 + r + 0.4) * EX
    return -pos_j, qi, qj


@pytest.mark.parametrize("r", [0.36, 0.4, 0.55, 0.69])
def test_hb_ideal_geometry_is_radial_morse(r):
    nl = dna2.make_neighbor_list()
    p, rc = first_struct(dna2.hydrogen_bonding(nl))
    dr, qi, qj = ideal_hb_pair(r)
    inside, F, Ti, Tj, E, _ = O.pair(_abi.TERM_HBOND, dr, qi, qj, rc, p)
    eps, a, r0 = dna2.HYDR_EPS, dna1.HYDR_A, dna1.HYDR_R0
    expect = eps * (1 - math.exp(-(r - r0) * a)) ** 2 - eps * dna1.hb_shift_factor()
    assert inside and E == pytest.approx(expect, rel=1e-12)
    # all angles at extrema: no torque; the force is the radial derivative along a1
    dE = 2 * eps * a * math.exp(-(r - r0) * a) * (1 - math.exp(-(r - r0) * a))
    assert np.allclose(Ti, 0, atol=1e-12) and np.allclose(Tj, 0, atol=1e-12)
    assert F == pytest.approx(dE * EX, abs=1e-12)  # force on i pulls it towards j when dE > 0


def test_hb_non_complementary_is_zero():
    nl = dna2.make_neighbor_list()
    p, rc = first_struct(dna2.hydrogen_bonding(nl), ("A", "G"))
    dr, qi, qj = ideal_hb_pair(0.4)
    _, F, Ti, Tj, E, _ = O.pair(_abi.TERM_HBOND, dr, qi, qj, rc, p)
    assert E == 0.0 and not np.any(F) and not np.any(Ti)


@pytest.mark.parametrize("alpha", [0.1, 0.4, 0.69, 0.9])
def test_hb_theta4_scan(alpha):
    """Rotating j about the base-base axis changes only theta4 (t0 = pi): E = f1(r0) f4(alpha)."""
    nl = dna2.make_neighbor_list()
    p, rc = first_struct(dna2.hydrogen_bonding(nl))
    dr, qi, qj = ideal_hb_pair(0.4)
    qj = H.quat_mul(H.delta_quat(EX, alpha), qj)
    _, _, _, _, E, _ = O.pair(_abi.TERM_HBOND, dr, qi, qj, rc, p)
    a, b, t0, ts, tc = dna1.HB_F4_THETA4
    f4 = 1 - a * alpha ** 2 if alpha < ts else b * (tc - alpha) ** 2
    e0 = -dna2.HYDR_EPS * dna1.hb_shift_factor()
    assert E == pytest.approx(e0 * f4, rel=1e-12)


def test_excluded_volume_single_site_pair_is_lj():
    """a1 of i along -x and of j along +x: only the BACK-BACK pair is within its cut-off."""
    nl = dna1.make_neighbor_list()
    p, rc = first_struct(dna1.excluded_volume(nl))
    qi, qj = H.axes_quat(-EX, EZ), H.axes_quat(EX, EZ)
    for s in (0.6, 0.65, 0.7):
        d = s + 0.8
        _, F, Ti, Tj, E, _ = O.pair(_abi.TERM_EXCVOL, -d * EX, qi, qj, rc, p)
        sigma, rstar, b, rcut = dna1.EXCL[1]
        if s < rstar:
            expect = 4 * 2.0 * ((sigma / s) ** 12 - (sigma / s) ** 6)
        else:
            expect = 2.0 * b * (s - rcut) ** 2
        assert E == pytest.approx(expect, rel=1e-12)
        assert np.allclose(Ti, 0, atol=1e-12)  # central pair along the site axis


def test_debye_screened_coulomb_and_charges():
    nl = dna2.make_neighbor_list()
    p, rc = first_struct(dna2.debye_huckel(nl))
    mk, pref, bs, rcd, rh = dna2.debye_params()
    qi, qj = H.axes_quat(-EX, EZ), H.axes_quat(EX, EZ)
    # grooved back site (-0.34, 0.3408, 0): with these frames the BACK-BACK vector is
    # d - 0.68 along x plus 0.6816 along y
    for d in (1.2, 2.0, 2.5):
        rb = math.hypot(d - 0.68, 0.0)
        _, _, _, _, E, _ = O.pair(_abi.TERM_DEBYE, -d * EX, qi, qj, rc, p, 1.0, 0.5)
        back_i = io.rotate(qi, dna2.BACK_SITE)[0]
        back_j = d * EX + io.rotate(qj, dna2.BACK_SITE)[0]
        r = np.linalg.norm(back_j - back_i)
        expect = pref * math.exp(mk * r) / r if r < rh else bs * (r - rcd) ** 2
        assert E == pytest.approx(0.5 * (expect if r < rcd else 0.0), rel=1e-12), (d, rb)
        _, _, _, _, E0, _ = O.pair(_abi.TERM_DEBYE, -d * EX, qi, qj, rc, p, 0.0, 1.0)
        assert E0 == 0.0


def test_stacking_ideal_geometry_is_radial_morse():
    """B directly above A with the same frame: theta4 = theta5 = theta6 = 0, cos(phi) = 0,
    so E_stack = f1(d) with the stacking epsilon; FENE closed form on the back sites."""
    f = dna1.bonded_force()
    P = f._pack(["backbone"])[0]
    q = H.axes_quat(EX, EZ)
    for d in (0.36, 0.4, 0.6, 0.74):  # base-base beyond its exc. vol. cut-off (0.335)
        r = O.bond(P, d * EZ, q, q, 0, 0)
        eps = dna1.STCK_BASE_EPS + dna1.STCK_FACT_EPS * dna1.kT_from_temperature(296.15)
        a, r0 = dna1.STCK_A, dna1.STCK_R0
        expect = eps * (1 - math.exp(-(d - r0) * a)) ** 2 - eps * dna1._stack_shift_factor()
        assert r["e_stack"] == pytest.approx(expect, rel=1e-12)
        dd = d - dna1.FENE_R0
        fene = -0.5 * 2.0 * math.log(1 - dd * dd / 0.0625) if abs(dd) < 0.25 else 1.0e12
        assert r["e_fene"] == pytest.approx(fene, rel=1e-12)
        assert r["e_bexc"] == 0.0


def test_broken_fene_bond_sentinel():
    f = dna1.bonded_force(stacking=False)
    P = f._pack(["backbone"])[0]
    q = H.axes_quat(EX, EZ)
    r = O.bond(P, (dna1.FENE_R0 + 0.26) * EZ, q, q)
    assert r["e_fene"] == 1.0e12 and not np.any(r["FB"])


@pytest.mark.parametrize("model,expect_t1", [(dna1, None), (dna2, None)])
def test_coax_stacked_geometry(model, expect_t1):
    """j directly above i with the same frame: theta4 = theta5 = 0, theta6 = pi,
    theta1 = pi, cos(phi3) = 0 -> E = f2(d) * f4-theta1 factor of the model."""
    nl = model.make_neighbor_list()
    p, rc = first_struct(model.coaxial_stacking(nl))
    q = H.axes_quat(EX, EZ)
    d = 0.4
    _, _, _, _, E, _ = O.pair(_abi.TERM_COAX, -d * EZ, q, q, rc, p)
    k, r0, rcf = p.f2_k, p.f2_r0, p.f2_rc
    f2 = 0.5 * k * ((d - r0) ** 2 - (rcf - r0) ** 2)
    a, b, t0, ts, tc = list(p.t1)
    dt = abs(math.pi - t0)
    f4 = 1 - a * dt * dt
    if p.t1_mode == 0:
        g1 = 2 * f4  # f4(pi) + f4(2 pi - pi)
    else:
        g1 = f4 + (p.t1_sa * (math.pi - p.t1_sb) ** 2 if math.pi > p.t1_sb else 0.0)
    assert E == pytest.approx(f2 * g1, rel=1e-12)


@pytest.mark.parametrize("t4_enabled,factor", [(True, 8.0), (False, 4.0)])
def test_cross_stacking_symmetrisation(t4_enabled, factor):
    """With f4 == 1 for every angle, the symmetrised t4 / t7 / t8 factors are 2 each."""
    nl = dna1.make_neighbor_list()
    cs = dna1.cross_stacking(nl)
    flat = (0.0, 0.0, 0.0, 10.0, 10.0)
    for key in cs.params.keys():
        v = dict(cs.params[key])
        v.update(f4_t1=flat, f4_t23=flat, f4_t4=flat, f4_t78=flat, t4_enabled=t4_enabled)
        cs.params[key] = v
    p, rc = first_struct(cs)
    rng = np.random.default_rng(3)
    for _ in range(5):
        qi, qj = H.random_quats(rng, 2)
        dr = rng.normal(size=3)
        base_i = io.rotate(qi, [0.4, 0, 0])[0]
        base_j = io.rotate(qj, [0.4, 0, 0])[0]
        target = 0.6 * dr / np.linalg.norm(dr)
        # choose dr so that the base-base distance is 0.6 (inside the f2 harmonic branch)
        d = -(target - base_j + base_i)
        _, _, _, _, E, _ = O.pair(_abi.TERM_CROSS, d, qi, qj, rc, p)
        f2 = 0.5 * 47.5 * ((0.6 - 0.575) ** 2 - (0.675 - 0.575) ** 2)
        assert E == pytest.approx(factor * f2, rel=1e-12)


# ------------------------------------------------------------------ finite differences
def _system_fd(system, force, n_check=4, seed=0, rtol=1e-4, atol=1e-6):
    orc = O.from_system(system, [force])
    mask = 1 << force._term
    ref = orc.compute(mask=mask)

    def energy(pos, quat):
        orc.set_state(pos, quat)
        return orc.compute(mask=mask)["E"].sum()

    F, T = ref["F"], ref["T"]
    assert np.abs(F).max() > 0, "no active pairs in the test system"
    fmag = np.linalg.norm(F, axis=1)
    idx = list(np.argsort(fmag)[::-1][:n_check])
    rng = np.random.default_rng(seed)
    idx += list(rng.choice(system.n, size=2, replace=False))
    for i in idx:
        f_fd, t_fd = H.fd_forces_torques(energy, system.pos.copy(), system.quat.copy(), int(i))
        assert np.allclose(F[i], f_fd, rtol=rtol, atol=atol * max(1, np.abs(F).max())), (i, F[i], f_fd)
        assert np.allclose(T[i], t_fd, rtol=rtol, atol=atol * max(1, np.abs(T).max())), (i, T[i], t_fd)
    orc.set_state(system.pos, system.quat)
    # Newton III and total angular momentum (about the origin, periodic images unwrapped
    # by using the minimum-image pair vectors implicitly: the test boxes are large)
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * max(1, np.abs(F).max())
    L = np.cross(system.pos, F).sum(axis=0) + T.sum(axis=0)
    assert np.abs(L).max() < 1e-8 * max(1, np.abs(F).max() * 10)


def _duplex(seed=0, n=12, dx=0.05, dq=0.08):
    rng = np.random.default_rng(seed)
    s = systems.duplex_system(systems.random_sequence(rng, n), box=60.0)
    return H.jiggle(s, seed=seed + 1, dx=dx, dq=dq)


def _coax_system():
    import census
    return H.jiggle(census.nicked_helix(n_bp=6, copies=3), seed=3, dx=0.02, dq=0.03)


FD_CASES = {
    "excvol_dna2": (lambda: _duplex(1, dx=0.12, dq=0.2), lambda nl: dna2.excluded_volume(nl)),
    "hbond_dna2": (lambda: _duplex(2), lambda nl: dna2.hydrogen_bonding(nl)),
    "cross_dna1": (lambda: _duplex(3), lambda nl: dna1.cross_stacking(nl)),
    "cross_rna2": (lambda: _duplex(4), lambda nl: rna2.cross_stacking(nl)),
    "hbondcross_dna2": (lambda: _duplex(5), lambda nl: dna2.hbond_cross(nl)),
    "hbondcross_rna2": (lambda: _duplex(6), lambda nl: rna2.hbond_cross(nl)),
    "coax_dna1": (_coax_system, lambda nl: dna1.coaxial_stacking(nl)),
    "coax_dna2": (_coax_system, lambda nl: dna2.coaxial_stacking(nl)),
    "coax_rna2": (_coax_system, lambda nl: rna2.coaxial_stacking(nl)),
    "debye_dna2": (lambda: _duplex(7), lambda nl: dna2.debye_huckel(nl)),
    "debye_rna2": (lambda: _duplex(8), lambda nl: rna2.debye_huckel(nl)),
    "bonded_dna1": (lambda: _duplex(9), lambda nl: dna1.bonded_force()),
    "bonded_dna2_seq": (lambda: _duplex(10), lambda nl: dna2.bonded_force(average=False)),
    "bonded_rna2": (lambda: _duplex(11), lambda nl: rna2.bonded_force()),
}


@pytest.mark.parametrize("case", sorted(FD_CASES))
def test_forces_torques_are_energy_gradients(case):
    make_sys, make_force = FD_CASES[case]
    system = make_sys()
    force = make_force(dna1.make_neighbor_list())
    _system_fd(system, force)


def test_coax_rna_and_dna_are_active_in_fd_systems():
    system = _coax_system()
    for model in (dna1, dna2, rna2):
        orc = O.from_system(system, [model.coaxial_stacking(model.make_neighbor_list())])
        assert orc.compute()["E"][_abi.E_CXSTACK] < 0


def test_fused_equals_separate():
    for model in (dna1, dna2, rna2):
        system = _duplex(12)
        nl = model.make_neighbor_list()
        cs = model.cross_stacking(nl) if model is rna2 else dna1.cross_stacking(nl)
        sep = O.from_system(system, [model.hydrogen_bonding(nl), cs]).compute()
        fus = O.from_system(system, [model.hbond_cross(nl)]).compute()
        assert np.allclose(sep["E"], fus["E"], rtol=1e-12, atol=1e-12)
        assert np.allclose(sep["F"], fus["F"], rtol=1e-10, atol=1e-12)
        assert np.allclose(sep["T"], fus["T"], rtol=1e-10, atol=1e-12)


def test_energy_invariant_under_global_rotation_and_translation():
    system = _duplex(13)
    forces, _ = dna2.forces()
    e0 = O.from_system(system, forces).compute()["E"]
    R = systems.haar_rotations(np.random.default_rng(1), 1)[0]
    qR = io.frames_to_quaternions(R[None])[0]
    pos = system.pos @ R.T + np.array([1.3, -2.1, 0.7])
    quat = np.array([H.quat_mul(qR, q) for q in system.quat])
    moved = systems.System(pos, quat, system.typeid, system.bonds, system.box, system.strand, system.bases,
                           system.n3, system.n5, system.charge)
    e1 = O.from_system(moved, forces).compute()["E"]
    assert np.allclose(e0, e1, rtol=1e-10, atol=1e-10)


def test_pair_torque_j_matches_swapped_evaluation():
    """Evaluating (j, i) gives j's force and torque: the evaluator's torque_j is the
    same physics seen from the other side (needed for half-list use)."""
    system = _duplex(14)
    nl = dna2.make_neighbor_list()
    for f in (dna2.excluded_volume(nl), dna2.hbond_cross(nl), dna2.coaxial_stacking(nl), dna2.debye_huckel(nl)):
        table, rcut = f._pack(["A", "G", "C", "T"])
        n_checked = 0
        for i in range(system.n):
            for j in range(i + 1, system.n):
                dr = system.pos[i] - system.pos[j]
                if np.linalg.norm(dr) > 2.5:
                    continue
                ti, tj = system.typeid[i], system.typeid[j]
                a = O.pair(f._term, dr, system.quat[i], system.quat[j], rcut[ti * 4 + tj], table[ti * 4 + tj])
                b = O.pair(f._term, -dr, system.quat[j], system.quat[i], rcut[tj * 4 + ti], table[tj * 4 + ti])
                assert a[4] + a[5] == pytest.approx(b[4] + b[5], rel=1e-10, abs=1e-12)
                assert np.allclose(a[1], -b[1], rtol=1e-8, atol=1e-10)
                assert np.allclose(a[3], b[2], rtol=1e-8, atol=1e-10)
                n_checked += 1
        assert n_checked > 10
