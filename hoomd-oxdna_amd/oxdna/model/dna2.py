"""oxDNA2: grooved backbone site, updated constants, Debye-Hueckel electrostatics.

Public API as reference oxdna/model/dna2.py. oxDNA2 changes against oxDNA1
(reference oxdna/model/dna2.py:24-55): a grooved BACK site used by FENE, excluded
volume and Debye-Hueckel (stacking and coaxial keep the ungrooved back reference),
FENE r0, stacking / HB strengths, a coaxial term with a harmonic theta1 and no
dihedral, and the Debye-Hueckel term with its salt-dependent derived constants.
"""

import itertools
import math

from ..forces import (OxDNABonded, ExcludedVolume, HydrogenBonding, HBondCross, CoaxialStacking,
                      DebyeHuckel)
from ..io import DNA_BASES
from . import dna1

POS_MM_BACK1, POS_MM_BACK2 = -0.3400, 0.3408
BACK_SITE = (POS_MM_BACK1, POS_MM_BACK2, 0.0)

FENE_R0 = 0.7564
STCK_BASE_EPS = 1.3523
STCK_FACT_EPS = 2.6717
HYDR_EPS = 1.0678

STCK_FACT_EPS_SEQ = 0.18
STCK_SEQ = {
    ("G", "C"): 1.69339, ("C", "G"): 1.74669, ("G", "G"): 1.61295, ("C", "C"): 1.61295,
    ("G", "A"): 1.59887, ("T", "C"): 1.59887, ("A", "G"): 1.61898, ("C", "T"): 1.61898,
    ("T", "G"): 1.66322, ("C", "A"): 1.66322, ("G", "T"): 1.68032, ("A", "C"): 1.68032,
    ("A", "T"): 1.56166, ("T", "A"): 1.64311, ("A", "A"): 1.84642, ("T", "T"): 1.58952,
}
HYDR_SEQ = {("A", "T"): 0.88537, ("T", "A"): 0.88537, ("C", "G"): 1.23238, ("G", "C"): 1.23238}

CXST_K = 58.5
CXST_F4_THETA1 = (2.0, 10.9032, math.pi - 0.25, 0.65, 0.769231)
CXST_THETA1_SA = 20.0
CXST_THETA1_SB = math.pi - 0.1 * (math.pi - (math.pi - 0.25))

DH_LAMBDA_FACTOR = 0.3616455
DH_PREFACTOR = 0.0543
DEFAULT_SALT = 0.5


def _pairs(types):
    return itertools.combinations_with_replacement(types, 2)


def stacking_eps_table(t_kelvin=296.15):
    """Sequence-averaged stacking epsilon as a uniform 4x4 table."""
    eps = STCK_BASE_EPS + STCK_FACT_EPS * dna1.kT_from_temperature(t_kelvin)
    return [[eps] * 4 for _ in range(4)]


def sequence_stacking_eps_table(t_kelvin=296.15):
    """Sequence-dependent stacking epsilon: STCK_<pair> * (1 - f + 9 f kT)."""
    scale = 1.0 - STCK_FACT_EPS_SEQ + 9.0 * STCK_FACT_EPS_SEQ * dna1.kT_from_temperature(t_kelvin)
    return [[STCK_SEQ[(bi, bj)] * scale for bj in DNA_BASES] for bi in DNA_BASES]


def _stacking_params(t_kelvin, average=True):
    sp = dna1.stacking_params(t_kelvin)
    table = stacking_eps_table(t_kelvin) if average else sequence_stacking_eps_table(t_kelvin)
    sp["stack_eps"] = tuple(table[i][j] for i in range(4) for j in range(4))
    return sp


def bonded_force(bond_type="backbone", temperature=296.15, stacking=True, average=True):
    force = OxDNABonded()
    params = dna1._backbone_params(FENE_R0, BACK_SITE)
    if stacking:
        params.update(_stacking_params(temperature, average=average))
    force.params[bond_type] = params
    return force


def excluded_volume(nlist, types=DNA_BASES, r_cut=dna1.EXCVOL_R_CUT):
    force = ExcludedVolume(nlist=nlist, default_r_cut=r_cut, mode="none")
    for key in _pairs(types):
        force.params[key] = dna1._excl_pair(BACK_SITE)
    return force


def hb_epsilon(t1, t2, average=True):
    if average:
        return HYDR_EPS if dna1.is_complementary(t1, t2) else 0.0
    return HYDR_SEQ.get((t1, t2), 0.0)


def hydrogen_bonding(nlist, types=DNA_BASES, average=True, r_cut=dna1.HB_R_CUT):
    force = HydrogenBonding(nlist=nlist, default_r_cut=r_cut, mode="none")
    for a, b in _pairs(types):
        force.params[(a, b)] = dict(eps=hb_epsilon(a, b, average), shift_factor=dna1.hb_shift_factor(),
                                    base_site=dna1.BASE_SITE, f1=dna1._hb_f1(),
                                    f4_t1=dna1.HB_F4_THETA1, f4_t23=dna1.HB_F4_THETA23,
                                    f4_t4=dna1.HB_F4_THETA4, f4_t78=dna1.HB_F4_THETA78)
    return force


def hbond_cross(nlist, types=DNA_BASES, average=True, r_cut=None):
    force = HBondCross(nlist=nlist,
                       default_r_cut=max(dna1.HB_R_CUT, dna1.CROSS_R_CUT) if r_cut is None else r_cut,
                       mode="none")
    for a, b in _pairs(types):
        force.params[(a, b)] = dict(
            base_site=dna1.BASE_SITE, eps=hb_epsilon(a, b, average),
            shift_factor=dna1.hb_shift_factor(), f1=dna1._hb_f1(),
            hb_f4_t1=dna1.HB_F4_THETA1, hb_f4_t23=dna1.HB_F4_THETA23,
            hb_f4_t4=dna1.HB_F4_THETA4, hb_f4_t78=dna1.HB_F4_THETA78,
            f2=dna1.CRST_F2, cs_f4_t1=dna1.CRST_F4_THETA1, cs_f4_t23=dna1.CRST_F4_THETA23,
            cs_f4_t4=dna1.CRST_F4_THETA4, cs_f4_t78=dna1.CRST_F4_THETA78)
    return force


def coaxial_stacking(nlist, types=DNA_BASES, r_cut=dna1.COAX_R_CUT):
    force = CoaxialStacking(nlist=nlist, default_r_cut=r_cut, mode="none")
    f2 = (CXST_K,) + tuple(dna1.CXST_F2[1:])
    for key in _pairs(types):
        force.params[key] = dict(stack_site=dna1.STACK_SITE, stack_back_ref=dna1.STACK_BACK_REF,
                                 gamma=dna1.GAMMA, f2=f2, f4_t1=CXST_F4_THETA1,
                                 f4_t4=dna1.CXST_F4_THETA4, f4_t56=dna1.CXST_F4_THETA56,
                                 f5_phi3=dna1.CXST_F5_PHI3, t1_mode=1, t1_sa=CXST_THETA1_SA,
                                 t1_sb=CXST_THETA1_SB, phi3_enabled=False)
    return force


def debye_from_lambda(lambda_factor, prefactor, t_kelvin, salt):
    """(minus_kappa, prefactor, b_smooth, rc, rhigh) of a screened-Coulomb term.

    lambda = lambda_factor sqrt(kT / 0.1) / sqrt(salt), rhigh = 3 lambda, and the
    quadratic tail B (r - rc)^2 is fixed by matching value and slope of
    prefactor e^{-r/lambda} / r at rhigh."""
    kT = dna1.kT_from_temperature(t_kelvin)
    lam = lambda_factor * math.sqrt(kT / 0.1) / math.sqrt(salt)
    rh = 3.0 * lam
    # slope match: with u = e^{-rh/lam} pref / rh, u' = -u (1/lam + 1/rh); B (rh-rc)^2 = u,
    # 2 B (rh - rc) = u'  =>  rh - rc = 2u/u', B = u'^2 / (4u).
    u = math.exp(-rh / lam) * prefactor / rh
    du = -u * (1.0 / lam + 1.0 / rh)
    rc = rh - 2.0 * u / du
    b_smooth = du * du / (4.0 * u)
    return -1.0 / lam, prefactor, b_smooth, rc, rh


def debye_params(t_kelvin=296.15, salt=DEFAULT_SALT):
    """Derived Debye-Hueckel parameters."""
    return debye_from_lambda(DH_LAMBDA_FACTOR, DH_PREFACTOR, t_kelvin, salt)


def debye_huckel(nlist, types=DNA_BASES, t_kelvin=296.15, salt=DEFAULT_SALT, r_cut=None):
    mk, pref, b_smooth, rc, rhigh = debye_params(t_kelvin, salt)
    if r_cut is None:
        r_cut = rc + 2.0 * math.sqrt(POS_MM_BACK1 ** 2 + POS_MM_BACK2 ** 2)
    force = DebyeHuckel(nlist=nlist, default_r_cut=r_cut, mode="none")
    for key in _pairs(types):
        force.params[key] = dict(back_site=BACK_SITE, debye=(mk, pref, b_smooth, rc, rhigh))
    return force


def make_neighbor_list(buffer=0.4):
    return dna1.make_neighbor_list(buffer=buffer)


def forces(nlist=None, temperature=296.15, salt=DEFAULT_SALT, fused=False, average=True):
    """The oxDNA2 force field: ``([bonded, excluded volume, hb, cross | hbond_cross,
    coaxial, debye], nlist)``."""
    if nlist is None:
        nlist = make_neighbor_list()
    if fused:
        base_base = [hbond_cross(nlist, average=average)]
    else:
        base_base = [hydrogen_bonding(nlist, average=average), dna1.cross_stacking(nlist)]
    return ([bonded_force(temperature=temperature, average=average), excluded_volume(nlist),
             *base_base, coaxial_stacking(nlist),
             debye_huckel(nlist, t_kelvin=temperature, salt=salt)], nlist)
