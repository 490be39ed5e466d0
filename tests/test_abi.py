"""The C ABI boundary without a GPU: struct layouts agree between C and the ctypes
mirrors, the library loads and exports every function include/oxdna_hip.h declares,
and entry points fail cleanly (status + message) when no device is present."""

import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from oxdna import _abi, _engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oxdna_hip.h")
STRUCTS = {
    "oxdna_excvol_params": _abi.ExcVolParams,
    "oxdna_hbond_params": _abi.HBondParams,
    "oxdna_crossstacking_params": _abi.CrossStackingParams,
    "oxdna_hbondcross_params": _abi.HBondCrossParams,
    "oxdna_coaxstacking_params": _abi.CoaxStackingParams,
    "oxdna_debye_params": _abi.DebyeParams,
    "oxdna_bonded_params": _abi.BondedParams,
    "oxdna_box": _abi.Box,
    "oxdna_tuning": _abi.Tuning,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oxdna_[a-z_0-9]+)\s*\(", text)))


def test_struct_layouts_match_c(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "oxdna_hip.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    for line in filter(None, out):
        cname, what, val = line.split()
        cls = STRUCTS[cname]
        if what == "sizeof":
            assert C.sizeof(cls) == int(val), cname
        else:
            assert getattr(cls, what).offset == int(val), (cname, what)


def _lib():
    if not os.path.exists(_engine.LIB_PATH):
        if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
            pytest.skip("engine library not built and hipcc unavailable")
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "hoomd-oxdna_amd", "csrc")])
    return _engine.load()


def test_library_exports_every_declared_function():
    lib = _lib()
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, f"declared in the header but not exported: {missing}"
    assert set(_engine.declared_symbols()) <= set(declared_functions())
    assert lib.oxdna_abi_version() == 1


def test_errors_cross_the_boundary_as_status_codes():
    lib = _lib()
    ctx = C.c_void_p()
    rc = lib.oxdna_ctx_create(0, None, C.byref(ctx))
    if rc == 0:  # a GPU is present: exercise an invalid call instead
        assert lib.oxdna_set_pair_params(ctx, 99, 4, None, None) != 0
        assert b"term" in lib.oxdna_last_error(ctx)
        lib.oxdna_ctx_destroy(ctx)
    else:
        assert rc < 0
        assert len(lib.oxdna_last_error(None)) > 0
    # NULL context is rejected everywhere
    assert lib.oxdna_compute(None, 1, 1) != 0
    assert lib.oxdna_run(None, 1, 1, 1) != 0


def test_drivers_reject_more_particles_than_row_indices_hold():
    """Neighbour indices keep 30 bits (the row keys' top bits hold the pair class): the
    stateless drivers reject N >= 2^30 on their arguments, before any device work (the
    device pointers below are never dereferenced)."""
    lib = _lib()
    fake = C.c_void_p(256)
    box = _abi.Box()
    box.L[0] = box.L[1] = box.L[2] = 10.0
    params = (_abi.ExcVolParams * 1)()
    rcut = (C.c_double * 1)(4.0)
    rc = lib.oxdna_pair_forces(_abi.TERM_EXCVOL, fake, fake, None, 0, 1 << 30, fake, fake, None, C.byref(box),
                               fake, fake, fake, 1, params, rcut, None, None)
    assert rc != 0 and b"2^30" in lib.oxdna_last_error(None)
    bparams = (_abi.BondedParams * 1)()
    rc = lib.oxdna_bonded_forces(fake, fake, None, 0, None, 1 << 30, fake, fake, C.byref(box), fake, fake, 1, fake,
                                 bparams, 1, None)
    assert rc != 0 and b"2^30" in lib.oxdna_last_error(None)


def test_simulation_without_gpu_fails_loudly():
    lib = _lib()
    ctx = C.c_void_p()
    if lib.oxdna_ctx_create(0, None, C.byref(ctx)) == 0:
        lib.oxdna_ctx_destroy(ctx)
        pytest.skip("a GPU is present")
    from oxdna import sim
    with pytest.raises(_engine.EngineError):
        sim.Simulation(device=sim.GPU())
    with pytest.raises(RuntimeError):
        sim.device.CPU()
