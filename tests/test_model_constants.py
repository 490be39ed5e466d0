"""Model layer: derived constants pinned by closed forms, the force-field assembly API
(reference oxdna/model/{dna1,dna2,rna2}.py), parameter schemas (reference
oxdna/forces.py) and the Watson-Crick table (reference pytest/test_hbond.py:27-30)."""

import math

import numpy as np
import pytest

from oxdna import forces as F
from oxdna.model import dna1, dna2, rna2


def test_kT_is_tenth_of_T_over_300K():
    assert dna1.kT_from_temperature(300.0) == pytest.approx(0.1)
    assert dna1.kT_from_temperature(296.15) == pytest.approx(0.0987166666666)


def test_morse_shift_factors_vanish_at_cutoff():
    """shift = (1 - exp(-(rc - r0) a))^2 makes eps (1 - e^{-(r-r0)a})^2 - eps*shift = 0 at rc."""
    for a, rc, r0, sf in ((dna1.HYDR_A, dna1.HYDR_RC, dna1.HYDR_R0, dna1.hb_shift_factor()),
                          (dna1.STCK_A, dna1.STCK_RC, dna1.STCK_R0, dna1._stack_shift_factor())):
        assert (1 - math.exp(-(rc - r0) * a)) ** 2 - sf == pytest.approx(0.0, abs=1e-15)


@pytest.mark.parametrize("model,salt", [(dna2, 0.5), (dna2, 0.1), (rna2, 1.0), (rna2, 0.3)])
def test_debye_smoothing_is_c1_and_matches_closed_form(model, salt):
    mk, q, B, rc, rh = model.debye_params(296.15, salt)
    lam = -1.0 / mk
    assert rh == pytest.approx(3 * lam)
    # value and slope of the screened Coulomb branch equal the quadratic tail at rhigh
    v = q * math.exp(-rh / lam) / rh
    dv = -v * (1 / lam + 1 / rh)
    assert B * (rh - rc) ** 2 == pytest.approx(v, rel=1e-12)
    assert 2 * B * (rh - rc) == pytest.approx(dv, rel=1e-12)
    # closed forms of oxDNA's DNA2Interaction::init (x = rhigh, l = lambda)
    x, l = rh, lam
    assert rc == pytest.approx(x * (q * x + 3 * q * l) / (q * (x + l)), rel=1e-12)
    assert B == pytest.approx(-(math.exp(-x / l) * q * q * (x + l) ** 2) / (-4 * x ** 3 * l * l * q), rel=1e-12)


def test_debye_lambda_scaling():
    """lambda = lambda_factor sqrt(kT / 0.1) / sqrt(salt): 1/sqrt(salt) and sqrt(T) laws."""
    mk1 = dna2.debye_params(300.0, 0.5)[0]
    mk4 = dna2.debye_params(300.0, 2.0)[0]
    assert mk4 / mk1 == pytest.approx(2.0)
    assert -1.0 / mk1 == pytest.approx(dna2.DH_LAMBDA_FACTOR / math.sqrt(0.5))


def test_cutoffs():
    assert dna1.HB_R_CUT == pytest.approx(0.783775 + 0.8)
    assert dna1.CROSS_R_CUT == pytest.approx(1.5)
    assert dna1.COAX_R_CUT == pytest.approx(0.6222222 + 0.68)
    nl = dna2.make_neighbor_list()
    dh = dna2.debye_huckel(nl)
    rc = dna2.debye_params()[3]
    assert dh.r_cut[("A", "T")] == pytest.approx(rc + 2 * math.hypot(0.34, 0.3408))


def test_watson_crick_complementarity():
    assert dna1.hb_epsilon("A", "T") > 0 and dna1.hb_epsilon("G", "C") > 0
    for a, b in [("A", "A"), ("A", "G"), ("A", "C"), ("G", "G"), ("G", "T"), ("C", "T")]:
        assert dna1.hb_epsilon(a, b) == 0.0
    # oxRNA2 sequence-dependent adds the G-U wobble (U shares T's type)
    assert rna2.hb_epsilon("G", "T", average=False) > 0
    assert rna2.hb_epsilon("G", "T", average=True) == 0.0


def test_sequence_tables():
    for model in (dna1, dna2):
        t = model.sequence_stacking_eps_table(296.15)
        assert len(t) == 4 and all(len(r) == 4 for r in t)
        kT = dna1.kT_from_temperature(296.15)
        scale = 1 - 0.18 + 9 * 0.18 * kT
        assert t[0][3] == pytest.approx(model.STCK_SEQ[("A", "T")] * scale)  # A (0) then T (3)
    t = rna2.sequence_stacking_eps_table(296.15)
    assert t[1][2] == pytest.approx(rna2.STCK_SEQ[("G", "C")] * (1 + dna1.kT_from_temperature(296.15) * 1.97561))


@pytest.mark.parametrize("model", [dna1, dna2, rna2])
@pytest.mark.parametrize("fused", [False, True])
def test_forces_assembly_order(model, fused):
    forces, nl = model.forces(fused=fused)
    kinds = [type(f) for f in forces]
    base = [F.HBondCross] if fused else [F.HydrogenBonding, F.CrossStacking]
    expect = [F.OxDNABonded, F.ExcludedVolume, *base, F.CoaxialStacking]
    if model is not dna1:
        expect.append(F.DebyeHuckel)
    assert kinds == expect
    assert all(f.nlist is nl for f in forces[1:])
    for f in forces[1:]:
        for a in "AGCT":
            for b in "AGCT":
                f.params[(a, b)]  # every type pair is set (unordered keys)
    forces[0].params["backbone"]


def test_rna_fused_drops_theta4_in_cross_part():
    nl = rna2.make_neighbor_list()
    assert rna2.hbond_cross(nl).params[("A", "T")]["cs_t4_enabled"] is False
    assert dna2.hbond_cross(nl).params[("A", "T")]["cs_t4_enabled"] is True


def test_parameter_schemas_match_reference_keys():
    """The per-force parameter keys are the reference's (oxdna/forces.py)."""
    nl = dna1.make_neighbor_list()
    assert set(dna1.excluded_volume(nl).params[("A", "A")]) == {
        "eps", "base_site", "back_site", "back_back", "base_base", "base_back", "back_base"}
    assert set(dna1.hydrogen_bonding(nl).params[("A", "T")]) == {
        "eps", "shift_factor", "base_site", "f1", "f4_t1", "f4_t23", "f4_t4", "f4_t78"}
    assert set(dna1.cross_stacking(nl).params[("A", "T")]) == {
        "base_site", "f2", "f4_t1", "f4_t23", "f4_t4", "f4_t78", "t4_enabled"}
    assert set(dna1.coaxial_stacking(nl).params[("A", "T")]) == {
        "stack_site", "stack_back_ref", "gamma", "f2", "f4_t1", "f4_t4", "f4_t56", "f5_phi3", "f5_phi4",
        "t1_mode", "t1_sa", "t1_sb", "phi3_enabled", "rna_coax"}
    assert set(dna2.debye_huckel(nl).params[("A", "T")]) == {"back_site", "debye"}
    keys = set(dna1.bonded_force().params["backbone"])
    assert {"epsilon", "r0", "delta", "pos_back", "pos_base", "excl_eps", "bexc_base_base", "bexc_base_back",
            "bexc_back_base", "stack_enabled", "stack_f1", "stack_eps", "stack_rna"} <= keys


def test_typeparam_semantics():
    nl = dna1.make_neighbor_list()
    f = dna1.excluded_volume(nl)
    assert f.params[("T", "A")] is f.params[("A", "T")] or f.params[("T", "A")] == f.params[("A", "T")]
    with pytest.raises(KeyError):
        f.params[("A", "T")] = dict(eps=1.0)  # missing required keys
    with pytest.raises(KeyError):
        f.params[("A", "T")] = dict(f.params[("A", "T")], bogus=1.0)
    with pytest.raises(ValueError):
        f.params[("A", "T")] = dict(f.params[("A", "T")], back_back=(1.0, 2.0))
    b = F.OxDNABonded()
    b.params["backbone"] = dict(epsilon=2.0, r0=0.75, delta=0.25, pos_back=(-0.4, 0, 0), pos_base=(0.4, 0, 0),
                                excl_eps=2.0, bexc_base_base=(0, 0, 0, 0), bexc_base_back=(0, 0, 0, 0),
                                bexc_back_base=(0, 0, 0, 0))
    assert b.params["backbone"]["stack_enabled"] is False  # reference default
    table = b._pack(["backbone"])
    assert table[0].delta2 == pytest.approx(0.0625) and table[0].stacking_enabled == 0


def test_pack_tables_are_row_major_and_symmetric():
    nl = dna2.make_neighbor_list()
    hb = dna2.hydrogen_bonding(nl, average=False)
    table, rcut = hb._pack(["A", "G", "C", "T"])
    eps = np.array([t.eps for t in table]).reshape(4, 4)
    assert np.allclose(eps, eps.T)
    assert eps[0, 3] == pytest.approx(dna2.HYDR_SEQ[("A", "T")]) and eps[1, 2] == pytest.approx(dna2.HYDR_SEQ[("G", "C")])
    assert eps[0, 0] == 0.0
    assert all(r == pytest.approx(dna1.HB_R_CUT) for r in rcut)
