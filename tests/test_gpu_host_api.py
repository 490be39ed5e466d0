"""GPU: the C++ host layer (include/oxdna_hip.hpp: OxDNABondedForceCompute and the
AnisoPotentialPairOxDNA* computes, one engine context each) driven by
tests/cpp/host_api_check, against the FP64 oracle. Tolerances as in test_gpu_parity."""

import os
import subprocess
import sys

import numpy as np
import pytest

import helpers as H
import hostapi
import oracle as O
from oxdna.model import dna1, dna2, rna2

pytestmark = pytest.mark.gpu

E_REL, E_PER_NT = 1e-4, 1e-5
F_REL, F_ABS = 2e-4, 2e-4


@pytest.mark.parametrize("model,fused", [(dna2, True), (dna2, False), (rna2, True), (dna1, False)])
def test_cpp_force_computes_match_oracle(tmp_path, model, fused):
    assert os.path.exists(hostapi.EXE), "tests/cpp/host_api_check is not built (build() builds it)"
    system = H.jiggle(H.cfg1_system(2), seed=12)
    forces, _ = model.forces(fused=fused)
    src, dst = tmp_path / "sys.bin", tmp_path / "out.bin"
    hostapi.write_system(str(src), system, forces)
    r = subprocess.run([hostapi.EXE, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    sys.stdout.flush()
    assert r.returncode == 0, r.stderr
    got = hostapi.read_output(str(dst), system.n)
    orc = O.from_system(system, forces)
    assert set(got) == {f._term for f in forces}
    for f in forces:
        e, F, T = got[f._term]
        ref = orc.compute(mask=1 << f._term)
        n = system.n
        if f._term == 6:
            assert np.all(np.abs(e[:3] - ref["E"][:3]) <= E_REL * np.abs(ref["E"][:3]) + E_PER_NT * n)
        else:
            assert abs(e[0] - ref["E"].sum()) <= E_REL * np.abs(ref["E"]).sum() + E_PER_NT * n
        assert np.abs(F[:, :3] - ref["F"]).max() <= F_REL * np.abs(ref["F"]).max() + F_ABS
        assert np.abs(T - ref["T"]).max() <= F_REL * max(np.abs(ref["T"]).max(), 1e-12) + F_ABS
        assert np.abs(F[:, 3] - ref["Eh"]).max() <= F_REL * max(1.0, np.abs(ref["Eh"]).max()) + F_ABS
