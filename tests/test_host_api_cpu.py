"""The C++ host layer (include/oxdna_hip.hpp) builds against the engine library and,
without a GPU, reports the engine's error through oxdna::Error (exit status 3)."""

import os
import shutil
import subprocess

import pytest

import helpers as H
import hostapi
from oxdna import _engine
from oxdna.model import dna2


def test_host_layer_compiles_and_fails_cleanly_without_gpu(tmp_path):
    if shutil.which("g++") is None or not os.path.exists(_engine.LIB_PATH):
        pytest.skip("g++ or the engine library missing")
    exe = hostapi.build()
    system = H.cfg1_system(0)
    forces, _ = dna2.forces(fused=True)
    src = tmp_path / "sys.bin"
    hostapi.write_system(str(src), system, forces)
    r = subprocess.run([exe, str(src), str(tmp_path / "out.bin")], capture_output=True, text=True)
    import ctypes as C
    lib = _engine.load()
    ctx = C.c_void_p()
    if lib.oxdna_ctx_create(0, None, C.byref(ctx)) == 0:
        lib.oxdna_ctx_destroy(ctx)
        pytest.skip("a GPU is present (the GPU test covers this path)")
    assert r.returncode == 3, r.stderr
    assert "oxdna_ctx_create" in r.stderr
