"""ctypes wrapper of the CPU oracle (liboxdna_oracle.so). TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker / CPU baseline. Parameter tables come from the product's force objects
(the same C ABI structs the engine receives).
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboxdna_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OracleSystem(C.Structure):
    _fields_ = [("n", C.c_uint32), ("pos", _dp), ("quat", _dp), ("type", _ip), ("charge", _dp),
                ("L", C.c_double * 3), ("nb", C.c_uint32), ("bonds", _ip), ("btype", _ip),
                ("ntypes", C.c_int32), ("pair_params", C.c_void_p * 6), ("rcut", _dp * 6),
                ("nbt", C.c_uint32), ("bond_params", C.c_void_p)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def _load(path):
    L = C.CDLL(path)
    L.oracle_compute.argtypes = [C.POINTER(OracleSystem), C.c_uint32, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.oracle_pairs.argtypes = [C.POINTER(OracleSystem), C.c_double, C.c_int, _ip, C.c_size_t,
                               C.POINTER(C.c_size_t)]
    L.oracle_pair.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_void_p, C.c_double, C.c_double, _dp]
    L.oracle_bond.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _dp]
    L.oracle_census.argtypes = [C.POINTER(OracleSystem), C.c_uint32, _dp]
    L.oracle_compute_list.argtypes = [C.POINTER(OracleSystem), C.c_uint32, _ip, C.c_size_t, C.c_int, _dp, _dp,
                                      _dp]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def baseline_lib(out_dir):
    """The oracle compiled for timing on this host (bench.py's cpu_baseline leg):
    g++ -O3 -march=native with OpenMP, into out_dir. Returns (library, compile command)."""
    path = os.path.join(out_dir, "liboxdna_oracle_native.so")
    cmd = ["g++", "-O3", "-march=native", "-fopenmp", "-std=c++17", "-fPIC", "-ffp-contract=off",
           "-I" + os.path.join(HERE, "..", "include"), "-shared", os.path.join(HERE, "oxdna_oracle.cpp"), "-o", path]
    subprocess.check_call(cmd)
    shown = cmd[:7] + ["-Iinclude", "-shared", "oracle/oxdna_oracle.cpp"]
    return _load(path), " ".join(shown)


def _arr(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class Oracle:
    """Holds one system (arrays + force parameter tables) for repeated evaluation."""

    def __init__(self, pos, quat, typeid, box, bonds=(), charge=None, forces=(), types=None,
                 bond_types=("backbone",), btype=None):
        self.pos = _arr(pos, np.float64).reshape(-1, 3)
        self.quat = _arr(quat, np.float64).reshape(-1, 4)
        self.type = _arr(typeid, np.int32)
        self.charge = _arr(np.ones(len(self.pos)) if charge is None else charge, np.float64)
        self.bonds = _arr(np.asarray(bonds, dtype=np.int32).reshape(-1, 2), np.int32)
        self.btype = _arr(np.zeros(len(self.bonds)) if btype is None else btype, np.int32)
        self.box = np.asarray(box, dtype=float)[:3]
        self.types = list(types) if types is not None else ["A", "G", "C", "T"]
        self._keep = []
        self.mask = 0
        s = OracleSystem()
        s.n = len(self.pos)
        s.ntypes = len(self.types)
        s.nb = len(self.bonds)
        for a in range(3):
            s.L[a] = self.box[a]
        self.sys = s
        for f in forces:
            self.add_force(f, bond_types)
        self._bind()

    def add_force(self, f, bond_types=("backbone",)):
        from oxdna.forces import OxDNABonded
        if isinstance(f, OxDNABonded):
            table = f._pack(list(bond_types))
            self._keep.append(table)
            self.sys.nbt = len(table)
            self.sys.bond_params = C.cast(table, C.c_void_p)
        else:
            table, rcut = f._pack(self.types)
            self._keep += [table, rcut]
            self.sys.pair_params[f._term] = C.cast(table, C.c_void_p)
            self.sys.rcut[f._term] = C.cast(rcut, _dp)
        self.mask |= 1 << f._term

    def _bind(self):
        s = self.sys
        s.pos = self.pos.ctypes.data_as(_dp)
        s.quat = self.quat.ctypes.data_as(_dp)
        s.type = self.type.ctypes.data_as(_ip)
        s.charge = self.charge.ctypes.data_as(_dp)
        s.bonds = self.bonds.ctypes.data_as(_ip)
        s.btype = self.btype.ctypes.data_as(_ip)

    def set_state(self, pos=None, quat=None):
        if pos is not None:
            self.pos = _arr(pos, np.float64).reshape(-1, 3)
        if quat is not None:
            self.quat = _arr(quat, np.float64).reshape(-1, 4)
        self._bind()

    def compute(self, mask=None, use_cells=True):
        """dict(F (n,3), T (n,3), Eh (n,), E (8,), V (n,6)) of the terms in mask."""
        n = len(self.pos)
        F, T, Eh, E, V = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(8), np.zeros((n, 6))
        rc = lib().oracle_compute(C.byref(self.sys), self.mask if mask is None else mask, 1 if use_cells else 0,
                                  F.ctypes.data_as(_dp), T.ctypes.data_as(_dp), Eh.ctypes.data_as(_dp),
                                  E.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
        assert rc == 0
        return dict(F=F, T=T, Eh=Eh, E=E, V=V)

    def compute_list(self, pairs, nthreads=1, mask=None, library=None):
        """Forces, torques and per-term energies over a given pair list (i < j) and the
        bonds on nthreads threads: dict(F, T, E)."""
        n = len(self.pos)
        F, T, E = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(8)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        L = library if library is not None else lib()
        rc = L.oracle_compute_list(C.byref(self.sys), self.mask if mask is None else mask,
                                   pairs.ctypes.data_as(_ip), len(pairs), int(nthreads),
                                   F.ctypes.data_as(_dp), T.ctypes.data_as(_dp), E.ctypes.data_as(_dp))
        assert rc == 0
        return dict(F=F, T=T, E=E)

    def pairs(self, r_list, use_cells=True):
        cap = max(16, len(self.pos) * 128)
        buf = np.empty((cap, 2), dtype=np.int32)
        cnt = C.c_size_t()
        rc = lib().oracle_pairs(C.byref(self.sys), r_list, 1 if use_cells else 0,
                                buf.ctypes.data_as(_ip), cap, C.byref(cnt))
        assert rc == 0, "pair buffer too small"
        return buf[:cnt.value].copy()

    def census(self, mask=None):
        out = np.zeros(7 * 6)
        lib().oracle_census(C.byref(self.sys), self.mask if mask is None else mask, out.ctypes.data_as(_dp))
        return out.reshape(7, 6)


def from_system(system, forces, **kw):
    """Oracle over an oxdna.systems.System with the given force objects."""
    return Oracle(system.pos, system.quat, system.typeid, system.box, bonds=system.bonds,
                  charge=system.charge, forces=forces, **kw)


def pair(term, dr, qi, qj, rcut, params, chg_i=1.0, chg_j=1.0):
    """One nonbonded pair -> (inside, F_i, T_i, T_j, E, E2)."""
    out = np.zeros(11)
    d = _arr(dr, np.float64)
    a = _arr(qi, np.float64)
    b = _arr(qj, np.float64)
    inside = lib().oracle_pair(term, d.ctypes.data_as(_dp), a.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                               float(rcut), C.cast(C.pointer(params), C.c_void_p), chg_i, chg_j,
                               out.ctypes.data_as(_dp))
    return bool(inside), out[0:3], out[3:6], out[6:9], out[9], out[10]


def bond(params, Rc, qA, qB, tA=0, tB=0):
    """One bond (A, B), Rc = r_B - r_A -> dict(FB, TA, TB, e_fene, e_bexc, e_stack, vir)."""
    out = np.zeros(18)
    r = _arr(Rc, np.float64)
    a = _arr(qA, np.float64)
    b = _arr(qB, np.float64)
    lib().oracle_bond(C.cast(C.pointer(params), C.c_void_p), r.ctypes.data_as(_dp), a.ctypes.data_as(_dp),
                      b.ctypes.data_as(_dp), int(tA), int(tB), out.ctypes.data_as(_dp))
    return dict(FB=out[0:3], TA=out[3:6], TB=out[6:9], e_fene=out[9], e_bexc=out[10], e_stack=out[11],
                vir=out[12:18])


def prim(which, x, params):
    """Building block values (see oracle_prim): returns (value, derivative-like)."""
    L = lib()
    if not hasattr(L, "_prim_set"):
        L.oracle_prim.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L._prim_set = True
    p = _arr(params, np.float64)
    out = np.zeros(2)
    L.oracle_prim(int(which), float(x), p.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out[0], out[1]
