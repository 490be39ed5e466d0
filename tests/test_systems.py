"""Synthetic systems, oxDNA file I/O and replica bookkeeping (CPU only).

Generators are checked through the oracle's energies: no broken backbone bond
(FENE finite), no excluded-volume overlap, paired stems hydrogen-bonded."""

import numpy as np
import pytest

import oracle as O
from oxdna import ensemble, io, systems
from oxdna.model import dna2, rna2


def _energies(system, model):
    forces, _ = model.forces(fused=True)
    return O.from_system(system, forces).compute()["E"] / system.n


def test_hairpin_ensemble_sizes_and_energies():
    reps = systems.hairpin_ensemble(16, seed=3)
    sizes = [r.n for r in reps]
    assert all(12 <= n <= 70 for n in sizes)
    assert len(set(sizes)) > 4  # uneven
    for r in reps:
        e = _energies(r, dna2)
        assert e[0] < 0.5            # FENE: every bond well inside its range
        assert e[1] < 1e-3 and e[3] < 1e-3  # no bonded / nonbonded excluded-volume overlap
        assert e[4] < -0.2           # the stem is hydrogen-bonded
        assert e[2] < -0.4           # and stacked
        assert np.all(np.abs(r.pos) < r.box / 2)


def test_hairpin_loop_backbone_spacing():
    """Loop backbone sites (oxDNA2 grooved site) are spaced by the requested step, the
    stem-loop junctions included."""
    p, a1, a3, bases, _, _, _ = systems.hairpin("ACGTAC", 7)
    a2 = np.cross(a3, a1)
    back = p - 0.34 * a1 + 0.3408 * a2
    d = np.linalg.norm(np.diff(back, axis=0), axis=1)
    loop = d[5:5 + 8]
    assert np.allclose(loop, 0.75, atol=1e-9)


def test_cloverleaf_layout():
    s, info = systems.cloverleaf(seed=5)
    assert s.n == 76
    assert 4 <= min(info["arm_stems"]) and max(info["arm_stems"]) <= 7
    assert 3 <= min(info["arm_loops"]) and max(info["arm_loops"]) <= 8
    comp = {"A": "U", "U": "A", "G": "C", "C": "G"}
    # acceptor stem: the first and last `acceptor` nt are complementary, antiparallel
    k = info["acceptor"]
    assert all(comp[s.bases[i]] == s.bases[s.n - 1 - i] for i in range(k))
    e = _energies(s, rna2)
    assert e[0] < 0.5 and e[1] < 1e-3 and e[3] < 1e-3
    assert e[4] < -0.15  # four stems hydrogen-bonded


def test_cloverleaf_ensemble_bias_centres():
    reps, ends, k, d0 = systems.cloverleaf_ensemble(8, seed=5, d_range=(2.0, 30.0))
    assert len(reps) == 8 and ends.shape == (8, 2)
    assert d0[0] == 2.0 and d0[-1] == 30.0 and np.all(np.diff(d0) > 0)
    assert all(np.array_equal(r.pos, reps[0].pos) for r in reps)
    b = ensemble.HarmonicBias(ends, k, d0)
    off = np.cumsum([0] + [r.n for r in reps])
    pos = np.concatenate([r.pos for r in reps])
    e = b.energy(pos, [r.box for r in reps], off)
    dist = np.linalg.norm(reps[0].pos[-1] - reps[0].pos[0])
    assert np.allclose(e, k * (dist - d0) ** 2)


def test_concatenate_snapshots_offsets_and_bonds():
    reps = systems.hairpin_ensemble(5, seed=1)
    snap, off, boxes = ensemble.concatenate_snapshots([r.snapshot() for r in reps])
    assert list(off) == list(np.cumsum([0] + [r.n for r in reps]))
    assert snap.particles.N == off[-1]
    g = snap.bonds.group
    assert len(g) == sum(len(r.bonds) for r in reps)
    for r, rep in enumerate(reps):  # every bond stays inside its replica
        inside = (g[:, 0] >= off[r]) & (g[:, 0] < off[r + 1])
        assert np.all((g[inside, 1] >= off[r]) & (g[inside, 1] < off[r + 1]))
        assert np.array_equal(g[inside] - off[r], rep.bonds)
    assert boxes.shape == (5, 3)


def test_replica_size_limit():
    big = systems.duplex_gas(14, 20, box=40.0, seed=4, min_dist=1.5)  # 560 nt
    with pytest.raises(ValueError):
        ensemble.concatenate_snapshots([big.snapshot()])


def test_top_conf_round_trip(tmp_path):
    s = systems.hairpin_ensemble(1, seed=2)[0]
    top, conf = tmp_path / "h.top", tmp_path / "h.conf"
    s.write(str(top), str(conf))
    snap = io.build_snapshot(str(top), str(conf))
    assert snap.particles.N == s.n
    # positions come back box-centred, orientations as the same frames
    assert np.allclose(snap.particles.position, s.pos, atol=1e-9)
    q1, q2 = snap.particles.orientation, s.quat
    assert np.allclose(np.abs(np.sum(q1 * q2, axis=1)), 1.0, atol=1e-9)  # q and -q are one rotation
    assert [tuple(b) for b in snap.bonds.group] == [tuple(b) for b in s.bonds]
    t = io.read_topology(str(top))
    assert "".join(t["base"]) == "".join(s.bases)


def test_angular_velocity_and_kinetic_energy_closed_form():
    """A particle rotated by phi about z whose body-frame angular momentum is L: the
    quaternion momentum is p = q (x) (0, 2 L) (q* (x) p = (0, 2 L_body)), the lab-frame
    angular velocity R_z(phi) (L / I) and the rotational energy sum L^2 / (2 I)."""
    from oxdna import io
    phi = 0.7
    q = np.array([[np.cos(phi / 2), 0.0, 0.0, np.sin(phi / 2)]])
    L = np.array([0.3, -0.2, 0.5])
    inertia = np.array([[2.0, 1.0, 4.0]])
    qw, qv = q[0, 0], q[0, 1:]
    lw, lv = 0.0, 2.0 * L
    p = np.concatenate([[qw * lw - qv @ lv], qw * lv + lw * qv + np.cross(qv, lv)])[None, :]
    assert np.allclose(io.body_angular_momentum(q, p), L, atol=1e-15)
    wb = L / inertia[0]
    Rz = np.array([[np.cos(phi), -np.sin(phi), 0.0], [np.sin(phi), np.cos(phi), 0.0], [0.0, 0.0, 1.0]])
    assert np.allclose(io.angular_velocity(q, p, inertia)[0], Rz @ wb, atol=1e-15)
    v = np.array([[0.1, 0.2, -0.3]])
    ke = io.kinetic_energy(np.array([3.0]), v, q, p, inertia)
    assert ke == pytest.approx(0.5 * 3.0 * 0.14 + 0.5 * np.sum(L * L / inertia[0]), rel=1e-14)


def test_conf_trajectory_arguments():
    from oxdna import sim
    with pytest.raises(ValueError):
        sim.ConfTrajectory("x.conf", 0)
    with pytest.raises(ValueError):
        sim.ConfTrajectory("x.conf", 10, mode="r")
    w = sim.ConfTrajectory("x.conf", 10)
    assert w.period == 10 and w.frames == 0
