"""ctypes mirrors of the POD structs and constants of include/oxdna_hip.h.

These are the parameter blocks that cross the C ABI. Field order and meaning follow
the reference's parameter structs (each class cites its counterpart); the layout is
checked against the C compiler's (sizeof / offsetof) by tests/test_abi_layout.py.
"""

import ctypes as C

D3 = C.c_double * 3
D4 = C.c_double * 4
D5 = C.c_double * 5

# Term ids (bit positions of a term mask).
TERM_EXCVOL = 0
TERM_HBOND = 1
TERM_CROSS = 2
TERM_COAX = 3
TERM_DEBYE = 4
TERM_HBONDCROSS = 5
TERM_BONDED = 6
N_TERMS = 7

# Energy slots in split_energy.dat column order (reference pytest/_helpers.py:27-30).
E_FENE, E_BEXC, E_STACK, E_NEXC, E_HB, E_CRSTACK, E_CXSTACK, E_DEBYE = range(8)
ENERGY_NAMES = ["fene", "b_exc", "stack", "n_exc", "hb", "cr_stack", "cx_stack", "debye"]


class ExcVolParams(C.Structure):
    """OxDNAExcVolParams (reference src/EvaluatorPairOxDNAExcVol.h:30-39)."""
    _fields_ = [("eps", C.c_double), ("base_site", D3), ("back_site", D3),
                ("sigma", D4), ("rstar", D4), ("b", D4), ("rc", D4)]


class HBondParams(C.Structure):
    """OxDNAHBondParams (reference src/EvaluatorPairOxDNAHBond.h:30-37)."""
    _fields_ = [("eps", C.c_double), ("shift_factor", C.c_double), ("base_site", D3),
                ("f1_a", C.c_double), ("f1_r0", C.c_double), ("f1_rlow", C.c_double),
                ("f1_rhigh", C.c_double), ("f1_rclow", C.c_double), ("f1_rchigh", C.c_double),
                ("f1_blow", C.c_double), ("f1_bhigh", C.c_double),
                ("t1", D5), ("t23", D5), ("t4", D5), ("t78", D5)]


class CrossStackingParams(C.Structure):
    """OxDNACrossStackingParams (reference src/EvaluatorPairOxDNACrossStacking.h:30-36)."""
    _fields_ = [("base_site", D3),
                ("f2_k", C.c_double), ("f2_r0", C.c_double), ("f2_rc", C.c_double),
                ("f2_rlow", C.c_double), ("f2_rhigh", C.c_double), ("f2_rclow", C.c_double),
                ("f2_rchigh", C.c_double), ("f2_blow", C.c_double), ("f2_bhigh", C.c_double),
                ("t1", D5), ("t23", D5), ("t4", D5), ("t78", D5),
                ("t4_enabled", C.c_int32), ("_pad", C.c_int32)]


class HBondCrossParams(C.Structure):
    """OxDNAHBondCrossParams (reference src/EvaluatorPairOxDNAHBondCross.h:37-47) plus
    cs_t4_enabled (absent in the reference; needed for a fused oxRNA path)."""
    _fields_ = [("base_site", D3), ("eps", C.c_double), ("shift_factor", C.c_double),
                ("f1_a", C.c_double), ("f1_r0", C.c_double), ("f1_rlow", C.c_double),
                ("f1_rhigh", C.c_double), ("f1_rclow", C.c_double), ("f1_rchigh", C.c_double),
                ("f1_blow", C.c_double), ("f1_bhigh", C.c_double),
                ("hb_t1", D5), ("hb_t23", D5), ("hb_t4", D5), ("hb_t78", D5),
                ("f2_k", C.c_double), ("f2_r0", C.c_double), ("f2_rc", C.c_double),
                ("f2_rlow", C.c_double), ("f2_rhigh", C.c_double), ("f2_rclow", C.c_double),
                ("f2_rchigh", C.c_double), ("f2_blow", C.c_double), ("f2_bhigh", C.c_double),
                ("cs_t1", D5), ("cs_t23", D5), ("cs_t4", D5), ("cs_t78", D5),
                ("cs_t4_enabled", C.c_int32), ("_pad", C.c_int32)]


class CoaxStackingParams(C.Structure):
    """OxDNACoaxStackingParams (reference src/EvaluatorPairOxDNACoaxStacking.h:32-47)."""
    _fields_ = [("stack_site", D3), ("stack_back_ref", D3), ("gamma", C.c_double),
                ("f2_k", C.c_double), ("f2_r0", C.c_double), ("f2_rc", C.c_double),
                ("f2_rlow", C.c_double), ("f2_rhigh", C.c_double), ("f2_rclow", C.c_double),
                ("f2_rchigh", C.c_double), ("f2_blow", C.c_double), ("f2_bhigh", C.c_double),
                ("t1", D5), ("t4", D5), ("t56", D5), ("phi3", D4), ("phi4", D4),
                ("t1_mode", C.c_int32), ("phi3_enabled", C.c_int32),
                ("t1_sa", C.c_double), ("t1_sb", C.c_double),
                ("rna_coax", C.c_int32), ("_pad", C.c_int32)]


class DebyeParams(C.Structure):
    """OxDNADebyeParams (reference src/EvaluatorPairOxDNADebye.h:28-31)."""
    _fields_ = [("back_site", D3), ("minus_kappa", C.c_double), ("prefactor", C.c_double),
                ("b_smooth", C.c_double), ("rc", C.c_double), ("rhigh", C.c_double)]


class BondedParams(C.Structure):
    """oxdna_bonded_params (reference src/OxDNABondedParams.h:35-73)."""
    _fields_ = [("eps_backbone", C.c_double), ("r0_backbone", C.c_double),
                ("delta_backbone", C.c_double), ("delta2", C.c_double),
                ("back_site", D3), ("base_site", D3), ("excl_eps", C.c_double),
                ("bx_sigma", D3), ("bx_rstar", D3), ("bx_b", D3), ("bx_rc", D3),
                ("stacking_enabled", C.c_int32), ("rna_stacking", C.c_int32),
                ("stack_site", D3), ("stack_back_ref", D3), ("st_gamma", C.c_double),
                ("st_a", C.c_double), ("st_r0", C.c_double), ("st_rlow", C.c_double),
                ("st_rhigh", C.c_double), ("st_rclow", C.c_double), ("st_rchigh", C.c_double),
                ("st_blow", C.c_double), ("st_bhigh", C.c_double), ("st_shift_factor", C.c_double),
                ("st_eps", C.c_double * 16), ("t4", D5), ("t5", D5), ("p1", D4),
                ("stack_3_site", D3), ("stack_5_site", D3), ("bbvector_3", D3), ("bbvector_5", D3),
                ("tB1", D5), ("tB2", D5)]


class Box(C.Structure):
    _fields_ = [("L", D3)]


class Tuning(C.Structure):
    """oxdna_tuning: the autotuner's launch parameters (no reference counterpart)."""
    _fields_ = [("lanes_per_particle", C.c_int), ("particles_per_workgroup", C.c_int),
                ("replica_workgroup", C.c_int), ("replica_records_per_pass", C.c_int),
                ("nlist_buffer", C.c_double)]


PARAM_STRUCT = {
    TERM_EXCVOL: ExcVolParams,
    TERM_HBOND: HBondParams,
    TERM_CROSS: CrossStackingParams,
    TERM_COAX: CoaxStackingParams,
    TERM_DEBYE: DebyeParams,
    TERM_HBONDCROSS: HBondCrossParams,
}
