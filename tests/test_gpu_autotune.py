"""GPU: the autotuner's neighbour-list skin and replica tile depth.

* The skin tuner times every candidate, keeps the fastest, sets it on the attached
  neighbour lists and restores the state: with the lanes per particle pinned, a run
  after tuning is bit-identical to a run of a fresh context given the chosen skin
  directly, and the result still matches the FP64 oracle (tolerances of
  test_gpu_parity: per-term energy 1e-4 relative + 1e-5 per nucleotide).
* The replica kernel's tuned workgroup size and records per pass are reported, and
  the ensemble's energies at the tuned shape match the oracle replica by replica.
"""

import sys

import numpy as np
import pytest

import oracle as O
from oxdna import ensemble, sim, systems
from oxdna.model import dna1, dna2

pytestmark = pytest.mark.gpu

E_REL, E_PER_NT = 1e-4, 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sim(system, buffer, lanes=16, seed=7):
    forces, nl = dna2.forces(fused=True)
    nl.buffer = buffer
    s = sim.Simulation(device=sim.GPU(), seed=seed, lanes_per_particle=lanes)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(
        dt=0.003, forces=forces, methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    return s, forces, nl


def test_skin_tuner_restores_state_and_keeps_the_fastest():
    system = systems.duplex_gas(40, 20, box=60.0, seed=2)
    _say(f"skin tuner: {system.n} nt, box 60")
    s, forces, nl = _sim(system, 0.4)
    s.run(0)
    before = s.state.get_snapshot().particles
    cand = [0.15, 0.3, 0.5, 0.7]
    rates = s.tune_nlist_buffer(cand, steps=600)
    _say("rates (steps/s):", {k: round(v) for k, v in rates.items()})
    assert sorted(rates) == sorted(round(c, 6) for c in cand)
    assert all(r > 0 for r in rates.values())
    chosen = s.tuning()["nlist_buffer"]
    assert chosen in cand
    assert rates[round(chosen, 6)] == max(rates.values())
    assert nl.buffer == chosen  # the attached neighbour list carries the choice
    assert s.timestep == 0 and s.stats()["step"] == 0
    after = s.state.get_snapshot().particles
    for name in ("position", "orientation", "velocity", "angmom"):
        np.testing.assert_array_equal(getattr(after, name), getattr(before, name), err_msg=name)

    # the trajectory continues as a fresh context's at the chosen skin, bit for bit
    s.run(300)
    ref, _, _ = _sim(system, chosen)
    ref.run(300)
    a, b = s.state.get_snapshot().particles, ref.state.get_snapshot().particles
    for name in ("position", "orientation", "velocity", "angmom"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)

    # and its forces still match the oracle at the state reached
    got = np.array(list(s.term_energies().values()))
    snap = s.state.get_snapshot()
    moved = systems.System(snap.particles.position.copy(), snap.particles.orientation.copy(), system.typeid,
                           system.bonds, system.box, system.strand, system.bases, system.n3, system.n5,
                           system.charge)
    want = O.from_system(moved, forces).compute()["E"]
    assert np.all(np.abs(got - want) <= E_REL * np.abs(want) + E_PER_NT * system.n), (got, want)


def test_skin_tuner_default_grid_and_errors():
    system = systems.duplex_gas(8, 20, box=40.0, seed=4)
    s, _, _ = _sim(system, 0.4, lanes=0)
    rates = s.tune_nlist_buffer(steps=200)
    assert sorted(rates) == [round(0.1 * k, 6) for k in range(1, 9)]
    t = s.tuning()
    _say("tuning:", t)
    assert any(abs(t["nlist_buffer"] - 0.1 * k) < 1e-12 for k in range(1, 9))
    assert t["lanes_per_particle"] in (4, 8, 16, 32)
    assert t["particles_per_workgroup"] * t["lanes_per_particle"] == 256
    with pytest.raises(Exception, match="candidate buffers"):
        s.tune_nlist_buffer([0.0, 0.3], steps=100)
    with pytest.raises(Exception, match="steps"):
        s.tune_nlist_buffer([0.3], steps=0)
    s.run(100)  # the context still runs after a refused call


def test_replica_tile_depth_is_tuned_and_matches_oracle():
    reps = systems.hairpin_ensemble(24, seed=3, box=30.0)
    forces, _ = dna2.forces(fused=True)
    ens = ensemble.ReplicaEnsemble(device=sim.GPU(), seed=5)
    ens.create_state_from_snapshots([r.snapshot() for r in reps])
    ens.operations.integrator = sim.Integrator(
        dt=0.003, forces=forces, methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    ens.run(50)
    t = ens.tuning()
    _say("replica tuning:", t)
    assert t["replica_workgroup"] in (64, 128, 256, 512)
    nmax = max(r.n for r in reps)
    assert t["replica_records_per_pass"] >= ((nmax + 7) // 8) * 8 + 64
    with pytest.raises(Exception, match="one-system"):
        ens.tune_nlist_buffer([0.3], steps=10)
    e = ens.replica_energies()
    for r in range(len(reps)):
        snap = ens.replica_snapshot(r)
        sysr = reps[r]
        moved = systems.System(snap.particles.position.copy(), snap.particles.orientation.copy(), sysr.typeid,
                               sysr.bonds, sysr.box, sysr.strand, sysr.bases, sysr.n3, sysr.n5, sysr.charge)
        want = O.from_system(moved, forces).compute()["E"]
        assert np.all(np.abs(e[r, :8] - want) <= E_REL * np.abs(want) + E_PER_NT * sysr.n), (r, e[r, :8], want)
