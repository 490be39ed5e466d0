"""The CPU baseline's evaluation path (oracle_compute_list: a prebuilt pair list, OpenMP
threads) gives what the oracle's own whole-system driver gives: the same forces, torques
and per-term energies, bit for bit on one thread, to rounding on several (the threads'
partial sums are added in another order)."""

import numpy as np

import oracle as O
from oxdna import systems
from oxdna.model import dna2


def test_list_evaluation_matches_oracle_compute(tmp_path):
    system = systems.duplex_gas(12, 20, box=40.0, seed=2)
    forces, nl = dna2.forces(fused=True)
    orc = O.from_system(system, forces)
    ref = orc.compute()
    r_list = max(float(f.r_cut[("A", "A")]) for f in forces[1:]) + nl.buffer
    pairs = orc.pairs(r_list)
    one = orc.compute_list(pairs, 1)
    assert np.array_equal(one["F"], ref["F"]) and np.array_equal(one["T"], ref["T"])
    assert np.array_equal(one["E"], ref["E"])
    lib, cmd = O.baseline_lib(str(tmp_path))
    assert "-O3" in cmd and "-march=native" in cmd and "-fopenmp" in cmd
    many = orc.compute_list(pairs, 4, library=lib)
    scale = np.abs(ref["F"]).max()
    assert np.abs(many["F"] - ref["F"]).max() <= 1e-11 * scale
    assert np.abs(many["T"] - ref["T"]).max() <= 1e-11 * np.abs(ref["T"]).max()
    assert np.allclose(many["E"], ref["E"], rtol=1e-12, atol=1e-12)
