"""ctypes binding of the HIP engine's C ABI (include/oxdna_hip.h, liboxdna_hip.so).

There is no CPU fallback: loading fails with an error that names the missing
library when it has not been built (``python __graft_entry__.py`` / ``build()``).
"""

import ctypes as C
import os

from . import _abi

_LIB = None
LIB_NAME = "liboxdna_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u32 = C.c_uint32
_u64 = C.c_uint64

_SIGS = {
    "oxdna_last_error": (C.c_char_p, [_p]),
    "oxdna_abi_version": (C.c_int, []),
    "oxdna_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
    "oxdna_ctx_create": (C.c_int, [C.c_int, _p, C.POINTER(_p)]),
    "oxdna_ctx_destroy": (C.c_int, [_p]),
    "oxdna_set_state": (C.c_int, [_p, _u32, _dp, _dp, _ip, _dp, C.POINTER(_abi.Box)]),
    "oxdna_set_momenta": (C.c_int, [_p, _dp, _dp]),
    "oxdna_set_bonds": (C.c_int, [_p, _u32, _ip, _ip]),
    "oxdna_set_pair_params": (C.c_int, [_p, C.c_int, C.c_int, _p, _dp]),
    "oxdna_set_bond_params": (C.c_int, [_p, _u32, C.POINTER(_abi.BondedParams)]),
    "oxdna_set_nlist": (C.c_int, [_p, C.c_double]),
    "oxdna_compute": (C.c_int, [_p, _u32, C.c_int]),
    "oxdna_get_term_energies": (C.c_int, [_p, _dp]),
    "oxdna_get_forces": (C.c_int, [_p, _dp, _dp, _dp]),
    "oxdna_get_nlist_pairs": (C.c_int, [_p, _ip, C.c_size_t, C.POINTER(C.c_size_t)]),
    "oxdna_set_langevin": (C.c_int, [_p, C.c_double, C.c_double, C.c_double, C.c_double, _u64]),
    "oxdna_run": (C.c_int, [_p, _u32, _u32, C.c_int]),
    "oxdna_get_state": (C.c_int, [_p, _dp, _dp, _dp, _dp]),
    "oxdna_synchronize": (C.c_int, [_p]),
    "oxdna_get_stats": (C.c_int, [_p, C.POINTER(_u64), C.POINTER(_u32), C.POINTER(_u64)]),
    "oxdna_set_lanes_per_particle": (C.c_int, [_p, C.c_int]),
    "oxdna_set_replica_tiling": (C.c_int, [_p, C.c_int, C.c_int]),
    "oxdna_get_tuning": (C.c_int, [_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oxdna_get_tuning_ex": (C.c_int, [_p, C.POINTER(_abi.Tuning)]),
    "oxdna_tune_nlist_buffer": (C.c_int, [_p, _u32, _dp, _u32, _u32, _dp, _dp]),
    "oxdna_time_forces": (C.c_int, [_p, _u32, _u32, C.POINTER(C.c_double)]),
    "oxdna_time_steps": (C.c_int, [_p, _u32, _u32, C.POINTER(C.c_double)]),
    "oxdna_count_evaluations": (C.c_int, [_p, _u32, C.POINTER(_u64)]),
    "oxdna_set_replicas": (C.c_int, [_p, _u32, C.POINTER(_u32), _dp]),
    "oxdna_set_replica_bias": (C.c_int, [_p, _ip, _dp, _dp]),
    "oxdna_get_replica_energies": (C.c_int, [_p, _dp]),
    "oxdna_pair_forces": (C.c_int, [C.c_int, _p, _p, _p, C.c_size_t, _u32, _p, _p, _p, C.POINTER(_abi.Box),
                                    _p, _p, _p, C.c_int, _p, _dp, _p, _p]),
    "oxdna_pair_forces_noalloc": (C.c_int, [C.c_int, _p, _p, _p, C.c_size_t, _u32, _p, _p, _p,
                                            C.POINTER(_abi.Box), _p, _p, _p, C.c_int, _p, _dp, _p, _p]),
    "oxdna_bonded_forces": (C.c_int, [_p, _p, _p, C.c_size_t, _p, _u32, _p, _p, C.POINTER(_abi.Box), _p, _p,
                                      _u32, _p, C.POINTER(_abi.BondedParams), _u32, _p]),
}


class EngineError(RuntimeError):
    pass


def declared_symbols():
    return list(_SIGS)


def load(path=None):
    """Load the HIP engine library once; raise EngineError if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or os.environ.get("OXDNA_HIP_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise EngineError(f"HIP engine library not found at {path}; build it first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(lib, ctx, rc, what):
    if rc != 0:
        msg = lib.oxdna_last_error(ctx)
        raise EngineError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
