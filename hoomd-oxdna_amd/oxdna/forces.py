"""oxDNA force objects: the reference's Python force API over the HIP engine.

Mirrors reference oxdna/forces.py: the same seven classes, the same parameter keys,
``len_keys`` and defaults, the same energy-on-read properties. Parameters are packed
into the C ABI structs of include/oxdna_hip.h (``_pack``) and evaluated by the
engine of the :class:`oxdna.sim.Simulation` the force is attached to.

Differences from the reference, all additive:
  * ``HBondCross`` accepts ``cs_t4_enabled`` (default True = the reference's
    behaviour) so a fused path exists for oxRNA2, whose cross-stacking drops theta4;
  * every force reports ``energy``, ``forces`` and ``torques`` (the host engine's
    Force properties the reference's tests read, pytest/_helpers.py:131-132).
"""

import ctypes as C
import itertools

import numpy as np

from . import _abi
from ._typeparam import TypeParameterDict

F3 = (float, float, float)
F4 = (float,) * 4
F5 = (float,) * 5


class Force:
    """Base of every force: attachment to a Simulation and the read-out properties."""

    _term = None
    _cpp_class_name = None

    def __init__(self):
        self._simulation = None

    def _require_attached(self):
        if self._simulation is None:
            raise RuntimeError(f"{type(self).__name__} is not attached to a simulation "
                               "(run the simulation first)")
        return self._simulation

    @property
    def energy(self):
        """Total energy of this force at the current state (computed on read)."""
        return self._require_attached()._force_energy(self)

    @property
    def forces(self):
        """Per-particle forces, (N, 3) float64."""
        return self._require_attached()._force_arrays(self)[0]

    @property
    def torques(self):
        """Per-particle lab-frame torques, (N, 3) float64."""
        return self._require_attached()._force_arrays(self)[1]

    @property
    def energies(self):
        """Per-particle energies (half of every pair / bond energy), (N,) float64."""
        return self._require_attached()._force_arrays(self)[2]

    @property
    def virials(self):
        """Per-particle virial (xx, xy, xz, yy, yz, zz), (N, 6) float64."""
        return self._require_attached()._force_arrays(self)[3]


class AnisotropicPair(Force):
    """Base of the nonbonded oxDNA pair forces (host engine's md.pair.aniso)."""

    _struct = None

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__()
        if mode != "none":
            raise ValueError("oxDNA pair forces support only mode='none' (no energy shift)")
        self.nlist = nlist
        self.mode = mode
        self.r_cut = TypeParameterDict(len_keys=2, scalar=float,
                                       default_value=None if default_r_cut is None else float(default_r_cut))
        self.params = None  # set by subclasses

    def _fill(self, s, v):  # pragma: no cover - subclass hook
        raise NotImplementedError

    def _pack(self, types):
        """ntypes^2 parameter structs and COM cut-offs (row-major [ti][tj])."""
        nt = len(types)
        self.params.require(types)
        table = (self._struct * (nt * nt))()
        rcut = (C.c_double * (nt * nt))()
        for i, a in enumerate(types):
            for j, b in enumerate(types):
                self._fill(table[i * nt + j], self.params[(a, b)])
                try:
                    rcut[i * nt + j] = float(self.r_cut[(a, b)])
                except (KeyError, TypeError):
                    raise KeyError(f"r_cut for ({a!r}, {b!r}) is not set and there is no default_r_cut")
        return table, rcut


def _set3(dst, v):
    dst[0], dst[1], dst[2] = v


def _setn(dst, v):
    for k, x in enumerate(v):
        dst[k] = x


class OxDNABonded(Force):
    """oxDNA bonded backbone force between 3'-5' neighbours (reference
    oxdna/forces.py:15-112): FENE + bonded excluded volume + (oxDNA or oxRNA)
    stacking, keyed by bond type."""

    _term = _abi.TERM_BONDED
    _cpp_class_name = "OxDNABondedForceCompute"

    def __init__(self):
        super().__init__()
        self.params = TypeParameterDict(
            dict(
                epsilon=float, r0=float, delta=float,
                pos_back=F3, pos_base=F3, excl_eps=float,
                bexc_base_base=F4, bexc_base_back=F4, bexc_back_base=F4,
                stack_enabled=bool, stack_site=F3, stack_back_ref=F3, stack_gamma=float,
                stack_f1=(float,) * 9, stack_eps=(float,) * 16,
                stack_f4_t4=F5, stack_f4_t5=F5, stack_f5_p1=F4,
                stack_rna=bool, stack_3_site=F3, stack_5_site=F3,
                stack_bbvector_3=F3, stack_bbvector_5=F3,
                stack_f4_tB1=F5, stack_f4_tB2=F5,
            ),
            len_keys=1,
            defaults={
                "stack_enabled": False, "stack_site": (0.0,) * 3, "stack_back_ref": (0.0,) * 3,
                "stack_gamma": 0.0, "stack_f1": (0.0,) * 9, "stack_eps": (0.0,) * 16,
                "stack_f4_t4": (0.0,) * 5, "stack_f4_t5": (0.0,) * 5, "stack_f5_p1": (0.0,) * 4,
                "stack_rna": False, "stack_3_site": (0.0,) * 3, "stack_5_site": (0.0,) * 3,
                "stack_bbvector_3": (0.0,) * 3, "stack_bbvector_5": (0.0,) * 3,
                "stack_f4_tB1": (0.0,) * 5, "stack_f4_tB2": (0.0,) * 5,
            },
        )

    @staticmethod
    def _fill(s, v):
        """dict -> oxdna_bonded_params (reference src/OxDNABondedParams.h:97-127, 183-220)."""
        C.memset(C.addressof(s), 0, C.sizeof(s))
        s.eps_backbone, s.r0_backbone, s.delta_backbone = v["epsilon"], v["r0"], v["delta"]
        s.delta2 = v["delta"] * v["delta"]
        _set3(s.back_site, v["pos_back"])
        _set3(s.base_site, v["pos_base"])
        s.excl_eps = v["excl_eps"]
        for k, key in enumerate(("bexc_base_base", "bexc_base_back", "bexc_back_base")):
            s.bx_sigma[k], s.bx_rstar[k], s.bx_b[k], s.bx_rc[k] = v[key]
        s.stacking_enabled = 1 if v["stack_enabled"] else 0
        if not v["stack_enabled"]:
            return
        _set3(s.stack_site, v["stack_site"])
        _set3(s.stack_back_ref, v["stack_back_ref"])
        s.st_gamma = v["stack_gamma"]
        (s.st_a, s.st_r0, s.st_rlow, s.st_rhigh, s.st_rclow, s.st_rchigh,
         s.st_blow, s.st_bhigh, s.st_shift_factor) = v["stack_f1"]
        _setn(s.st_eps, v["stack_eps"])
        _setn(s.t4, v["stack_f4_t4"])
        _setn(s.t5, v["stack_f4_t5"])
        _setn(s.p1, v["stack_f5_p1"])
        s.rna_stacking = 1 if v["stack_rna"] else 0
        if v["stack_rna"]:
            _set3(s.stack_3_site, v["stack_3_site"])
            _set3(s.stack_5_site, v["stack_5_site"])
            _set3(s.bbvector_3, v["stack_bbvector_3"])
            _set3(s.bbvector_5, v["stack_bbvector_5"])
            _setn(s.tB1, v["stack_f4_tB1"])
            _setn(s.tB2, v["stack_f4_tB2"])

    def _pack(self, bond_types):
        self.params.require(bond_types)
        table = (_abi.BondedParams * len(bond_types))()
        for k, name in enumerate(bond_types):
            self._fill(table[k], self.params[name])
        return table

    def getParams(self, bond_type):
        """Parameters of one bond type as a dict (reference OxDNABondedForceCompute::getParams)."""
        return dict(self.params[bond_type])

    @property
    def fene_energy(self):
        """Total FENE backbone energy (computed on read)."""
        return self._require_attached()._term_energies(self)[_abi.E_FENE]

    @property
    def bonded_excvol_energy(self):
        """Total bonded excluded-volume energy (computed on read)."""
        return self._require_attached()._term_energies(self)[_abi.E_BEXC]

    @property
    def stacking_energy(self):
        """Total stacking energy (computed on read)."""
        return self._require_attached()._term_energies(self)[_abi.E_STACK]


class ExcludedVolume(AnisotropicPair):
    """Nonbonded excluded volume (reference oxdna/forces.py:115-145)."""

    _term = _abi.TERM_EXCVOL
    _struct = _abi.ExcVolParams
    _cpp_class_name = "AnisoPotentialPairOxDNAExcVol"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(
            dict(eps=float, base_site=F3, back_site=F3,
                 back_back=F4, base_base=F4, base_back=F4, back_base=F4),
            len_keys=2)

    @staticmethod
    def _fill(s, v):
        s.eps = v["eps"]
        _set3(s.base_site, v["base_site"])
        _set3(s.back_site, v["back_site"])
        for k, key in enumerate(("back_back", "base_base", "base_back", "back_base")):
            s.sigma[k], s.rstar[k], s.b[k], s.rc[k] = v[key]


class HydrogenBonding(AnisotropicPair):
    """Watson-Crick hydrogen bonding (reference oxdna/forces.py:148-176)."""

    _term = _abi.TERM_HBOND
    _struct = _abi.HBondParams
    _cpp_class_name = "AnisoPotentialPairOxDNAHBond"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(
            dict(eps=float, shift_factor=float, base_site=F3, f1=(float,) * 8,
                 f4_t1=F5, f4_t23=F5, f4_t4=F5, f4_t78=F5),
            len_keys=2)

    @staticmethod
    def _fill(s, v):
        s.eps, s.shift_factor = v["eps"], v["shift_factor"]
        _set3(s.base_site, v["base_site"])
        (s.f1_a, s.f1_r0, s.f1_rlow, s.f1_rhigh, s.f1_rclow, s.f1_rchigh,
         s.f1_blow, s.f1_bhigh) = v["f1"]
        _setn(s.t1, v["f4_t1"]); _setn(s.t23, v["f4_t23"])
        _setn(s.t4, v["f4_t4"]); _setn(s.t78, v["f4_t78"])


class CrossStacking(AnisotropicPair):
    """Cross-stacking (reference oxdna/forces.py:179-207)."""

    _term = _abi.TERM_CROSS
    _struct = _abi.CrossStackingParams
    _cpp_class_name = "AnisoPotentialPairOxDNACrossStacking"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(
            dict(base_site=F3, f2=(float,) * 9, f4_t1=F5, f4_t23=F5, f4_t4=F5, f4_t78=F5,
                 t4_enabled=bool),
            len_keys=2, defaults={"t4_enabled": True})

    @staticmethod
    def _fill(s, v):
        _set3(s.base_site, v["base_site"])
        (s.f2_k, s.f2_r0, s.f2_rc, s.f2_rlow, s.f2_rhigh, s.f2_rclow, s.f2_rchigh,
         s.f2_blow, s.f2_bhigh) = v["f2"]
        _setn(s.t1, v["f4_t1"]); _setn(s.t23, v["f4_t23"])
        _setn(s.t4, v["f4_t4"]); _setn(s.t78, v["f4_t78"])
        s.t4_enabled = 1 if v["t4_enabled"] else 0


class HBondCross(AnisotropicPair):
    """Fused hydrogen bonding + cross-stacking sharing one BASE-BASE geometry
    (reference oxdna/forces.py:210-247). ``energy`` is the sum of both parts, like
    the reference; ``hb_energy`` / ``cross_energy`` give them apart."""

    _term = _abi.TERM_HBONDCROSS
    _struct = _abi.HBondCrossParams
    _cpp_class_name = "AnisoPotentialPairOxDNAHBondCross"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(
            dict(base_site=F3, eps=float, shift_factor=float, f1=(float,) * 8,
                 hb_f4_t1=F5, hb_f4_t23=F5, hb_f4_t4=F5, hb_f4_t78=F5,
                 f2=(float,) * 9, cs_f4_t1=F5, cs_f4_t23=F5, cs_f4_t4=F5, cs_f4_t78=F5,
                 cs_t4_enabled=bool),
            len_keys=2, defaults={"cs_t4_enabled": True})

    @staticmethod
    def _fill(s, v):
        _set3(s.base_site, v["base_site"])
        s.eps, s.shift_factor = v["eps"], v["shift_factor"]
        (s.f1_a, s.f1_r0, s.f1_rlow, s.f1_rhigh, s.f1_rclow, s.f1_rchigh,
         s.f1_blow, s.f1_bhigh) = v["f1"]
        _setn(s.hb_t1, v["hb_f4_t1"]); _setn(s.hb_t23, v["hb_f4_t23"])
        _setn(s.hb_t4, v["hb_f4_t4"]); _setn(s.hb_t78, v["hb_f4_t78"])
        (s.f2_k, s.f2_r0, s.f2_rc, s.f2_rlow, s.f2_rhigh, s.f2_rclow, s.f2_rchigh,
         s.f2_blow, s.f2_bhigh) = v["f2"]
        _setn(s.cs_t1, v["cs_f4_t1"]); _setn(s.cs_t23, v["cs_f4_t23"])
        _setn(s.cs_t4, v["cs_f4_t4"]); _setn(s.cs_t78, v["cs_f4_t78"])
        s.cs_t4_enabled = 1 if v["cs_t4_enabled"] else 0

    @property
    def hb_energy(self):
        return self._require_attached()._term_energies(self)[_abi.E_HB]

    @property
    def cross_energy(self):
        return self._require_attached()._term_energies(self)[_abi.E_CRSTACK]


class CoaxialStacking(AnisotropicPair):
    """Coaxial stacking, oxDNA1 / oxDNA2 / oxRNA forms (reference oxdna/forces.py:250-286)."""

    _term = _abi.TERM_COAX
    _struct = _abi.CoaxStackingParams
    _cpp_class_name = "AnisoPotentialPairOxDNACoaxStacking"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(
            dict(stack_site=F3, stack_back_ref=F3, gamma=float, f2=(float,) * 9,
                 f4_t1=F5, f4_t4=F5, f4_t56=F5, f5_phi3=F4, f5_phi4=F4,
                 t1_mode=int, t1_sa=float, t1_sb=float, phi3_enabled=bool, rna_coax=bool),
            len_keys=2, defaults={"f5_phi4": (0.0,) * 4, "rna_coax": False})

    @staticmethod
    def _fill(s, v):
        _set3(s.stack_site, v["stack_site"])
        _set3(s.stack_back_ref, v["stack_back_ref"])
        s.gamma = v["gamma"]
        (s.f2_k, s.f2_r0, s.f2_rc, s.f2_rlow, s.f2_rhigh, s.f2_rclow, s.f2_rchigh,
         s.f2_blow, s.f2_bhigh) = v["f2"]
        _setn(s.t1, v["f4_t1"]); _setn(s.t4, v["f4_t4"]); _setn(s.t56, v["f4_t56"])
        _setn(s.phi3, v["f5_phi3"]); _setn(s.phi4, v["f5_phi4"])
        if v["t1_mode"] not in (0, 1):
            raise ValueError("t1_mode must be 0 (oxDNA1 reflection) or 1 (oxDNA2 harmonic)")
        s.t1_mode = v["t1_mode"]
        s.t1_sa, s.t1_sb = v["t1_sa"], v["t1_sb"]
        s.phi3_enabled = 1 if v["phi3_enabled"] else 0
        s.rna_coax = 1 if v["rna_coax"] else 0


class DebyeHuckel(AnisotropicPair):
    """oxDNA2 / oxRNA2 Debye-Hueckel electrostatics (reference oxdna/forces.py:289-311)."""

    _term = _abi.TERM_DEBYE
    _struct = _abi.DebyeParams
    _cpp_class_name = "AnisoPotentialPairOxDNADebye"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        self.params = TypeParameterDict(dict(back_site=F3, debye=F5), len_keys=2)

    @staticmethod
    def _fill(s, v):
        _set3(s.back_site, v["back_site"])
        s.minus_kappa, s.prefactor, s.b_smooth, s.rc, s.rhigh = v["debye"]


PAIR_CLASSES = (ExcludedVolume, HydrogenBonding, CrossStacking, HBondCross, CoaxialStacking,
                DebyeHuckel)


def pack_pair_table(force, types):
    """(struct table, rcut table) of a pair force for the given type names."""
    return force._pack(list(types))
