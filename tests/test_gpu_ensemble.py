"""GPU parity of the batched replica path (oxdna_set_replicas, one persistent launch
for all replicas) against the FP64 CPU oracle, replica by replica, and against
one-system contexts run with the same random streams.

Tolerances are the one-system ones (test_gpu_parity): per-term energy
|dE| <= 1e-4 |E_ref| + 1e-5 N, per-particle force / torque
|d| <= 2e-4 max|ref| + 2e-4. The bias is checked against its closed form
k (|r_e - r_s| - d0)^2 and its analytic force (FP64 on the host).
"""

import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import ensemble, sim, systems
from oxdna.model import dna1, dna2, rna2

pytestmark = pytest.mark.gpu

E_REL, E_PER_NT = 1e-4, 1e-5
F_REL, F_ABS = 2e-4, 2e-4
KT = dna1.kT_from_temperature(296.15)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ensemble(reps, forces, seed=3, langevin=False):
    ens = ensemble.ReplicaEnsemble(device=sim.GPU(), seed=seed)
    ens.create_state_from_snapshots([r.snapshot() for r in reps])
    methods = [sim.Langevin(kT=KT)] if langevin else []
    ens.operations.integrator = sim.Integrator(dt=0.003, forces=forces, methods=methods)
    return ens


def _check_replicas(reps, forces, label, bias=None, tiling=None):
    ens = _ensemble(reps, forces)
    if tiling is not None:
        ens.set_tiling(*tiling)
    if bias is not None:
        ens.bias = bias
    ens.run(0)
    got_e = ens.replica_energies()
    _, f4, t3, _ = ens._eval(ens._mask)
    off = ens.offsets
    problems = []
    worst = [0.0, 0.0, 0.0]
    for r, rep in enumerate(reps):
        ref = O.from_system(rep, forces).compute()
        F_ref, T_ref = ref["F"].copy(), ref["T"].copy()
        e_ref = np.zeros(9)
        e_ref[:8] = ref["E"]
        if bias is not None:
            s, e = bias.ends[r]
            d = rep.pos[e] - rep.pos[s]
            d -= rep.box * np.rint(d / rep.box)
            dist = np.linalg.norm(d)
            e_ref[8] = bias.k[r] * (dist - bias.d0[r]) ** 2
            fe = -2.0 * bias.k[r] * (dist - bias.d0[r]) * d / dist
            F_ref[e] += fe
            F_ref[s] -= fe
        n = rep.n
        dE = np.abs(got_e[r] - e_ref)
        tolE = E_REL * np.abs(e_ref) + E_PER_NT * n
        sl = slice(int(off[r]), int(off[r + 1]))
        dF = np.abs(f4[sl, :3] - F_ref).max()
        dT = np.abs(t3[sl] - T_ref).max()
        fmax, tmax = np.abs(F_ref).max(), np.abs(T_ref).max()
        worst = [max(worst[0], (dE / np.maximum(np.abs(e_ref), 1e-300)).max(where=e_ref != 0, initial=0.0)),
                 max(worst[1], dF / fmax), max(worst[2], dT / tmax)]
        if not np.all(dE <= tolE):
            problems.append(f"replica {r} (n={n}): energies {got_e[r]} vs {e_ref}")
        if not dF <= F_REL * fmax + F_ABS:
            problems.append(f"replica {r}: force error {dF} (max {fmax})")
        if not dT <= F_REL * tmax + F_ABS:
            problems.append(f"replica {r}: torque error {dT} (max {tmax})")
    _say(f"[{label}] {len(reps)} replicas: worst rel dE={worst[0]:.1e}, dF/max={worst[1]:.1e}, "
         f"dT/max={worst[2]:.1e}")
    assert not problems, f"{label}: " + "; ".join(problems[:5])
    return ens


@pytest.mark.parametrize("fused", [True, False])
def test_hairpin_replicas_match_oracle(fused):
    reps = [H.jiggle(h, seed=10 + k, dx=0.03, dq=0.05) for k, h in enumerate(systems.hairpin_ensemble(12, seed=3))]
    forces, _ = dna2.forces(fused=fused)
    _check_replicas(reps, forces, f"hairpins dna2 fused={fused}")


def test_oxdna1_duplex_replicas_match_oracle():
    reps = [H.jiggle(H.cfg1_system(k), seed=20 + k) for k in range(5)]
    forces, _ = dna1.forces(fused=True)
    _check_replicas(reps, forces, "duplexes dna1")


def test_rna_cloverleaf_replicas_with_bias_match_oracle():
    reps, ends, k, d0 = systems.cloverleaf_ensemble(6, seed=5, d_range=(2.0, 30.0), k_bias=0.5)
    reps = [H.jiggle(r, seed=40 + i, dx=0.02, dq=0.03) for i, r in enumerate(reps)]
    forces, _ = rna2.forces(fused=True)
    bias = ensemble.HarmonicBias(ends, k, d0)
    ens = _check_replicas(reps, forces, "cloverleaf rna2 + bias", bias=bias)
    # bias energies are large for the far centres and vanish nowhere
    e = ens.replica_energies()[:, 8]
    assert np.all(e > 0) and e[-1] > e[0]


def test_uneven_sizes_one_launch_up_to_496():
    """Replicas from 8 to 496 nt in one launch (workgroup of 512 threads): 496 nt is the
    largest replica whose LDS layout (state, frames, bond and pair tables, the minimum
    of ncap + 64 pair records) fits the 160 KiB of one workgroup; a 497-nt replica is
    rejected with the limit in the message."""
    rng = np.random.default_rng(2)
    big = systems.duplex_gas(12, 20, box=40.0, seed=4, min_dist=1.5)  # 480 nt
    single = systems.duplex_system(systems.random_sequence(rng, 4), box=20.0)
    mid = H.jiggle(H.cfg1_system(6), seed=7)
    reps = [single, big, mid]
    forces, _ = dna2.forces(fused=True)
    _check_replicas(reps, forces, "uneven sizes")
    # the largest accepted replica (31 duplexes of 8 bp: 496 nt) against the oracle
    top = systems.duplex_gas(31, 8, box=40.0, seed=5, min_dist=1.5)
    _check_replicas([top, single], forces, "496-nt replica")
    # a larger one is refused (83 duplexes of 3 bp: 498 nt)
    over = systems.duplex_gas(83, 3, box=60.0, seed=6, min_dist=1.5)
    ens2 = ensemble.ReplicaEnsemble(device=sim.GPU(), seed=3)
    with pytest.raises(Exception, match="at most 496"):
        ens2.create_state_from_snapshots([over.snapshot()])


def _records(rep, forces):
    """Pair records of a replica: pairs whose FP64 minimum-image centre distance lies
    inside the largest cut-off of any term (the kernel's test)."""
    rmax = max(max(f._pack(["A", "G", "C", "T"])[1]) for f in forces[1:])
    d = rep.pos[:, None] - rep.pos[None]
    d -= rep.box * np.rint(d / rep.box)
    r2 = (d * d).sum(-1)[np.triu_indices(rep.n, 1)]
    return int((r2 < rmax * rmax).sum())


@pytest.mark.parametrize("tiling", [(128, 4), (256, 2), (512, 1)])
def test_pinned_tilings_match_oracle(tiling):
    """Every pinned replica-kernel shape gives the oracle's forces: (128, 4) holds fewer
    records per pass than the largest hairpin has, so its pairs run in several passes
    and the second ends are found through the incoming-record table with the pass
    check; the others run the same replicas in one pass."""
    reps = [H.jiggle(h, seed=60 + k, dx=0.03, dq=0.05) for k, h in enumerate(systems.hairpin_ensemble(12, seed=3))]
    forces, _ = dna2.forces(fused=True)
    ens = _check_replicas(reps, forces, f"hairpins tiling {tiling}", tiling=tiling)
    t = ens.tuning()
    most = max(_records(r, forces) for r in reps)
    _say(f"tiling {tiling}: {t['replica_records_per_pass']} records per pass, largest replica {most} records")
    assert t["replica_workgroup"] == tiling[0]
    if tiling == (128, 4):
        assert t["replica_records_per_pass"] < most, "the pinned depth should force several passes"


def test_multipass_small_replicas_are_bit_identical_and_follow_one_pass():
    """With several passes per step on replicas small enough for the incoming-record
    table, two runs are bit-identical, and 100 Langevin steps follow the one-pass shape
    up to FP32 summation order."""
    reps = list(systems.hairpin_ensemble(16, seed=23))
    forces, _ = dna2.forces(fused=True)
    out = []
    for tiling in ((128, 4), (128, 4), (512, 1)):
        ens = _ensemble(reps, forces, seed=9, langevin=True)
        ens.set_tiling(*tiling)
        ens.run(100)
        snap = ens.state.get_snapshot()
        out.append((snap.particles.position.copy(), snap.particles.velocity.copy()))
        del ens
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    dx = np.abs(out[0][0] - out[2][0]).max()
    dv = np.abs(out[0][1] - out[2][1]).max()
    _say(f"multi-pass vs one-pass after 100 steps: |dx| {dx:.2e}, |dv| {dv:.2e}")
    assert dx < 1e-3 and dv < 1e-2


def test_pinned_tiling_rejects_bad_shapes():
    reps = list(systems.hairpin_ensemble(4, seed=2))
    forces, _ = dna2.forces(fused=True)
    ens = _ensemble(reps, forces)
    with pytest.raises(Exception, match="workgroup"):
        ens.set_tiling(96, 1)
    with pytest.raises(Exception, match="depth_divisor"):
        ens.set_tiling(128, 3)
    big = systems.duplex_gas(4, 20, box=40.0, seed=4, min_dist=1.5)  # 160 nt
    ens2 = _ensemble([big], forces, langevin=True)
    ens2.set_tiling(128, 1)
    with pytest.raises(Exception, match="smaller than the largest replica"):
        ens2.run(0)
    ens2.set_tiling(0)  # back to the autotuner, which times the shapes at the first run
    ens2.run(1)
    assert ens2.tuning()["replica_workgroup"] >= 256


def test_replicas_follow_one_system_runs():
    """Replica r of an ensemble with seed s draws the random stream of a one-system run
    with seed s + r: after 50 Langevin steps both paths agree up to FP32 summation
    order (the pair sums run in different orders)."""
    base = systems.hairpin_ensemble(3, seed=8)
    forces, _ = dna2.forces(fused=True)
    ens = _ensemble(base, forces, seed=100, langevin=True)
    ens.run(50)
    snap = ens.state.get_snapshot()
    off = ens.offsets
    for r, rep in enumerate(base):
        f2, _ = dna2.forces(fused=True)
        s = sim.Simulation(device=sim.GPU(), seed=100 + r)
        s.create_state_from_snapshot(rep.snapshot())
        s.operations.integrator = sim.Integrator(dt=0.003, forces=f2, methods=[sim.Langevin(kT=KT)])
        s.run(50)
        one = s.state.get_snapshot()
        sl = slice(int(off[r]), int(off[r + 1]))
        dx = np.abs(snap.particles.position[sl] - one.particles.position).max()
        dv = np.abs(snap.particles.velocity[sl] - one.particles.velocity).max()
        moved = np.abs(one.particles.position - rep.pos).max()
        _say(f"replica {r} (n={rep.n}): |dx|={dx:.2e} |dv|={dv:.2e} (moved {moved:.3f})")
        assert dx < 1e-3 and dv < 1e-2
        assert moved > 1e-3


def test_replica_run_is_stable_and_reports_energies():
    reps = systems.hairpin_ensemble(32, seed=9)
    forces, _ = dna2.forces(fused=True)
    ens = _ensemble(reps, forces, seed=1, langevin=True)
    ens.run(3000)
    e = ens.replica_energies()
    pos = ens.state.get_snapshot().particles.position
    hb = e[:, 4] / np.diff(ens.offsets)
    _say(f"after 3000 steps: mean HB/nt {hb.mean():.4f}, min {hb.min():.4f}, fene max {e[:, 0].max():.3f}")
    assert np.isfinite(pos).all() and np.isfinite(e).all()
    assert np.all(e[:, 0] < 1e6)          # no broken backbone bond
    assert hb.mean() < -0.1               # stems stay paired on average


def test_replica_runs_are_bit_identical():
    """Every force, torque and energy sum of the replica kernel has a fixed order, so two
    runs from the same state and seed give identical bits: hairpins of 12-70 nt (one
    pass of pair records) and a 480-nt replica whose pairs need several passes."""
    rng = np.random.default_rng(5)
    reps = list(systems.hairpin_ensemble(24, seed=21))
    reps.append(systems.duplex_gas(12, 20, box=40.0, seed=4, min_dist=1.5))
    reps.append(systems.duplex_system(systems.random_sequence(rng, 6), box=20.0))
    forces, _ = dna2.forces(fused=True)
    out = []
    for trial in range(2):
        ens = _ensemble(reps, forces, seed=7, langevin=True)
        ens.run(400)
        snap = ens.state.get_snapshot()
        ens.run(0)
        _, f4, t3, _ = ens._eval(ens._mask)
        out.append((snap.particles.position.copy(), snap.particles.orientation.copy(),
                    snap.particles.velocity.copy(), f4.copy(), t3.copy(), ens.replica_energies().copy()))
        del ens
    names = ("position", "orientation", "velocity", "force", "torque", "replica energies")
    for name, a, b in zip(names, out[0], out[1]):
        assert np.array_equal(a, b), f"{name} differs between identical runs (max {np.abs(a - b).max()})"
    moved = np.abs(out[0][0] - np.concatenate([r.pos for r in reps])).max()
    _say(f"two 400-step runs of {len(reps)} replicas are bit-identical (moved up to {moved:.3f})")
    assert moved > 1e-3
