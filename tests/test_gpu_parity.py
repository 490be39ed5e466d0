"""GPU parity: the HIP engine (through the C ABI) against the FP64 CPU oracle.

Tolerances (FP32 engine vs FP64 oracle, stated per check):
  * per-term energy: |dE| <= 1e-4 |E_ref| + 1e-5 N  (1e-4 relative, plus a 1e-5 per
    nucleotide floor for terms that are ~0; the reference's own bar is 1e-4 per nt);
  * per-particle force / torque: |d| <= 2e-4 max|ref| + 2e-4 per component (FP32
    accumulation of O(30) pair terms whose largest are the stiff excluded-volume cores);
  * neighbour pair sets: bit-exact.
"""

import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import sim, systems
from oxdna.model import dna1, dna2, rna2

pytestmark = pytest.mark.gpu

E_REL, E_PER_NT = 1e-4, 1e-5
F_REL, F_ABS = 2e-4, 2e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine_eval(system, forces, lanes=0):
    """Attach forces to a Simulation, run(0), read per-force outputs."""
    s = sim.Simulation(device=sim.GPU(), lanes_per_particle=lanes)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces)
    s.run(0)
    out = {}
    for f in forces:
        e = s._term_energies(f)
        F, T, Eh, V = f.forces, f.torques, f.energies, f.virials
        out[f._term] = dict(E=e, F=F, T=T, Eh=Eh, V=V)
    return s, out


def check_against_oracle(system, forces, label, lanes=0):
    orc = O.from_system(system, forces)
    _, got = engine_eval(system, forces, lanes)
    n = system.n
    worst = {}
    problems = []
    for f in forces:
        ref = orc.compute(mask=1 << f._term)
        g = got[f._term]
        dE = np.abs(g["E"] - ref["E"])
        tolE = E_REL * np.abs(ref["E"]) + E_PER_NT * n
        fmax = max(np.abs(ref["F"]).max(), 1e-12)
        tmax = max(np.abs(ref["T"]).max(), 1e-12)
        dF = np.abs(g["F"] - ref["F"]).max()
        dT = np.abs(g["T"] - ref["T"]).max()
        dEh = np.abs(g["Eh"] - ref["Eh"]).max()
        dV = np.abs(g["V"] - ref["V"]).max()
        vmax = max(np.abs(ref["V"]).max(), 1e-12)
        name = type(f).__name__
        slots = [k for k in range(8) if ref["E"][k] != 0.0 or g["E"][k] != 0.0]
        erel = max([dE[k] / max(abs(ref["E"][k]), 1e-300) for k in slots], default=0.0)
        _say(f"[{label}] {name}: E/nt={np.round(ref['E'][slots] / n, 6).tolist()} rel dE={erel:.1e} "
             f"dF={dF:.1e} (max|F| {fmax:.1e}) dT={dT:.1e} (max|T| {tmax:.1e}) dEh={dEh:.1e} "
             f"dV={dV:.1e} (max|V| {vmax:.1e})")
        if not np.all(dE <= tolE):
            problems.append(f"{name}: energy {g['E']} vs {ref['E']}")
        if not dF <= F_REL * fmax + F_ABS:
            problems.append(f"{name}: force error {dF}")
        if not dT <= F_REL * tmax + F_ABS:
            problems.append(f"{name}: torque error {dT}")
        if not dEh <= F_REL * max(np.abs(ref['Eh']).max(), 1.0) + F_ABS:
            problems.append(f"{name}: per-particle energy error {dEh}")
        if not dV <= F_REL * vmax + F_ABS:
            problems.append(f"{name}: virial error {dV}")
        worst[name] = (dE.max(), dF, dT)
    assert not problems, f"{label}: " + "; ".join(problems)
    return worst


def test_device_is_gfx950():
    from oxdna import _engine
    lib = _engine.load()
    buf = C.create_string_buffer(64)
    assert lib.oxdna_device_name(0, buf, 64) == 0
    _say("device:", buf.value.decode())
    assert buf.value.decode().startswith("gfx950")


@pytest.mark.parametrize("model,fused", [(dna2, False), (dna2, True), (dna1, False), (dna1, True),
                                         (rna2, False), (rna2, True)])
def test_duplex_all_terms(model, fused):
    system = H.jiggle(H.cfg1_system(0), seed=2, dx=0.04, dq=0.06)
    forces, _ = model.forces(fused=fused)
    check_against_oracle(system, forces, f"{model.__name__.split('.')[-1]} fused={fused}")


def test_sequence_dependent_oxdna2():
    system = H.jiggle(H.cfg1_system(5), seed=3)
    forces, _ = dna2.forces(average=False)
    check_against_oracle(system, forces, "dna2 seq-dep")


def test_half_charged_ends_debye():
    rng = np.random.default_rng(7)
    system = systems.duplex_system(systems.random_sequence(rng, 12), box=20.0, half_charged_ends=True)
    system = H.jiggle(system, seed=4)
    nl = dna2.make_neighbor_list()
    check_against_oracle(system, [dna2.debye_huckel(nl)], "debye half-charged")


def test_gas_of_duplexes_parity():
    system = H.jiggle(systems.duplex_gas(24, 10, box=18.0, seed=11, min_dist=1.2), seed=5, dx=0.03, dq=0.05)
    forces, _ = dna2.forces(fused=True)
    check_against_oracle(system, forces, "gas dna2 fused")


@pytest.mark.parametrize("lanes", [4, 8, 16, 32])
def test_particles_per_workgroup_variants(lanes):
    """Every force-kernel layout (256 / lanes particles per workgroup: how the near pairs,
    bond slots and far pairs of a workgroup pack onto wavefronts) against the oracle."""
    system = H.jiggle(systems.duplex_gas(40, 12, box=30.0, seed=12, min_dist=1.2), seed=9, dx=0.03, dq=0.05)
    forces, _ = dna2.forces(fused=True)
    check_against_oracle(system, forces, f"gas dna2 fused, {lanes} lanes", lanes=lanes)


def test_large_grid_paths():
    """A grid that fills the GPU (4800 nt at 32 lanes per particle: more than 2048
    wavefronts) takes the other kernel layout (one far row entry per thread) and, in a
    run, a force launch plus the separate integrator kernel per step: forces against the
    oracle before and after a 300-step Langevin run on that path."""
    system = H.jiggle(systems.duplex_gas(200, 12, box=100.0, seed=21), seed=4, dx=0.01, dq=0.02)
    assert system.n * 32 // 64 > 2048
    forces, _ = dna2.forces(fused=True)
    check_against_oracle(system, forces, "large grid, 32 lanes", lanes=32)
    s = _langevin_sim(system, forces, seed=3, lanes=32)
    s.run(300)
    snap = s.state.get_snapshot()
    assert np.isfinite(snap.particles.position).all()
    assert forces[0].fene_energy / system.n < 1.0  # no backbone bond broke
    hot = systems.System(snap.particles.position, snap.particles.orientation, system.typeid, system.bonds,
                         system.box, system.strand, system.bases, system.n3, system.n5, system.charge)
    check_against_oracle(hot, dna2.forces(fused=True)[0], "large grid after 300 steps, 32 lanes", lanes=32)


def test_rows_cover_every_pair_after_a_hot_run():
    """A hot Langevin run of a duplex gas (many in-kernel row rebuilds and cell crossings:
    the run step bins each particle by its position before the force kick, so now and
    then a cell holds a hole) ends with rows that hold every pair closer than the
    largest cut-off (less a margin for the last step's motion) at the current positions."""
    system = H.jiggle(systems.duplex_gas(40, 12, box=30.0, seed=31), seed=5, dx=0.01, dq=0.02)
    forces, nl = dna2.forces(fused=True)
    s = _langevin_sim(system, forces, seed=11, kT=0.2)
    s.run(3000)
    st = s.stats()
    pos = s.state.get_snapshot().particles.position
    got = {tuple(p) for p in s.neighbor_pairs()}
    rcut = max(float(f.r_cut[("A", "A")]) for f in forces[1:])
    L = np.asarray(system.box, float)[:3]
    bonded = {tuple(sorted(b)) for b in np.asarray(system.bonds)}
    missing = []
    for i in range(system.n):
        d = pos[i + 1:] - pos[i]
        d -= L * np.round(d / L)
        for k in np.nonzero((d * d).sum(1) < (rcut - 0.05) ** 2)[0]:
            pair = (i, i + 1 + int(k))
            if pair not in bonded and pair not in got:
                missing.append(pair)
    _say(f"hot run: {st['nlist_builds']} row builds, {len(got)} row pairs, missing {len(missing)}")
    assert st["nlist_builds"] > 5
    assert not missing


def test_pinned_lanes_reproducible():
    """Same lanes per particle, same seed: two runs give bit-identical trajectories."""
    system = H.cfg1_system(2)
    out = []
    for _ in range(2):
        forces, _ = dna2.forces(fused=True)
        s = _langevin_sim(system, forces, seed=4, lanes=8)
        s.run(200)
        out.append(s.state.get_snapshot().particles.position.copy())
    assert np.array_equal(out[0], out[1])


def test_coaxial_stacking_nicked_helix():
    """A continuous helix of 3 duplexes with nicks: coaxial stacking across the nicks."""
    import census
    system = H.jiggle(census.nicked_helix(n_bp=8, copies=3), seed=8, dx=0.02, dq=0.03)
    for model in (dna1, dna2, rna2):
        forces, _ = model.forces(fused=True)
        worst = check_against_oracle(system, forces, f"nicked helix {model.__name__.split('.')[-1]}")
        assert "CoaxialStacking" in worst
    orc = O.from_system(system, [dna2.coaxial_stacking(dna2.make_neighbor_list())])
    assert orc.compute()["E"][6] < 0  # the coaxial term is active


def test_nlist_pairs_bit_exact():
    system = systems.duplex_gas(30, 10, box=20.0, seed=13, min_dist=1.0)
    forces, nl = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU())
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces)
    s.run(0)
    got = s.neighbor_pairs()
    rlist = max(float(f.r_cut[("A", "A")]) for f in forces[1:]) + nl.buffer
    orc = O.from_system(system, forces)
    ref = orc.pairs(rlist, use_cells=False)
    _say(f"nlist pairs: engine {len(got)} oracle {len(ref)} (r_list={rlist})")
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_unfused_compute_matches_fused():
    system = H.jiggle(H.cfg1_system(1), seed=6)
    forces, _ = dna2.forces(fused=False)
    s = sim.Simulation(device=sim.GPU())
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces)
    s.run(0)
    lib, ctx = s._lib, s._ctx
    e1, e2 = np.zeros(8), np.zeros(8)
    f1, f2 = np.zeros((system.n, 4)), np.zeros((system.n, 4))
    dp = C.POINTER(C.c_double)
    assert lib.oxdna_compute(ctx, s._mask, 1) == 0
    lib.oxdna_get_term_energies(ctx, e1.ctypes.data_as(dp))
    lib.oxdna_get_forces(ctx, f1.ctypes.data_as(dp), None, None)
    assert lib.oxdna_compute(ctx, s._mask, 0) == 0
    lib.oxdna_get_term_energies(ctx, e2.ctypes.data_as(dp))
    lib.oxdna_get_forces(ctx, f2.ctypes.data_as(dp), None, None)
    _say("fused vs unfused dE", np.abs(e1 - e2).max(), "dF", np.abs(f1 - f2).max())
    assert np.allclose(e1, e2, rtol=1e-6, atol=1e-6)
    assert np.allclose(f1, f2, rtol=1e-5, atol=1e-5)


def _langevin_sim(system, forces, seed=5, use_graph=True, kT=None, lanes=0):
    s = sim.Simulation(device=sim.GPU(), seed=seed, use_graph=use_graph, lanes_per_particle=lanes)
    s.create_state_from_snapshot(system.snapshot())
    kT = dna1.kT_from_temperature(296.15) if kT is None else kT
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces, methods=[sim.Langevin(kT=kT)])
    return s


def test_duplex_stays_bound_langevin():
    """Reference stability check (pytest/test_oxdna2.py:72-80): HB/nt < -0.15 after 2000 steps."""
    system = H.cfg1_system(3)
    forces, _ = dna2.forces()
    s = _langevin_sim(system, forces)
    s.run(2000)
    snap = s.state.get_snapshot()
    hb = forces[2].energy / system.n
    _say(f"after 2000 Langevin steps: hb/nt = {hb:.4f}, stats {s.stats()}")
    assert np.isfinite(snap.particles.position).all()
    assert hb < -0.15


def test_graph_and_direct_steps_agree():
    """With the lanes per particle pinned (the summation order), graph replays and direct
    steps give bit-identical trajectories."""
    system = H.cfg1_system(4)
    out = []
    for g in (True, False):
        forces, _ = dna2.forces(fused=True)
        s = _langevin_sim(system, forces, seed=9, use_graph=g, lanes=16)
        s.run(100)
        out.append(s.state.get_snapshot().particles.position.copy())
    _say("graph vs direct max |dx|", np.abs(out[0] - out[1]).max())
    assert np.array_equal(out[0], out[1])


def _tiny(pos, quat, typeid, bonds, strand, box=20.0):
    n = len(pos)
    n3, n5 = -np.ones(n, dtype=int), -np.ones(n, dtype=int)
    for a_, b_ in bonds:
        n5[a_], n3[b_] = b_, a_
    return systems.System(np.asarray(pos, float), np.asarray(quat, float), np.asarray(typeid), np.asarray(bonds),
                          [box] * 3, np.asarray(strand), ["ACGT"[t] for t in typeid], n3, n5)


def test_tiny_systems():
    """Edge cases: one nucleotide (all forces and energies zero; a run moves it and stays
    finite), a bonded dimer (bonded terms against the oracle; the bond partners are no
    nonbonded pair) and two free nucleotides beyond every cut-off (all zero)."""
    q0 = [1.0, 0.0, 0.0, 0.0]
    one = _tiny([[0.0, 0.0, 0.0]], [q0], [0], np.zeros((0, 2), int), [0])
    forces, _ = dna2.forces(fused=True)
    s, got = engine_eval(one, forces)
    for t, g in got.items():
        assert np.all(g["E"] == 0.0) and np.all(g["F"] == 0.0) and np.all(g["T"] == 0.0), t
    forces, _ = dna2.forces(fused=True)
    s = _langevin_sim(one, forces, seed=2)
    s.run(50)
    p = s.state.get_snapshot().particles.position
    assert np.isfinite(p).all() and np.abs(p).max() > 0.0

    ang = 2 * np.pi / 10.34
    dimer = _tiny([[0.0, 0.0, 0.0], [0.0, 0.0, 0.39]],
                  [q0, [np.cos(ang / 2), 0.0, 0.0, np.sin(ang / 2)]], [0, 2], [[0, 1]], [0, 0])
    for model in (dna1, dna2, rna2):  # (the oxRNA2 backbone is stretched past its FENE limit here:
        # the reference's broken-bond energy, 1e12 per bond, and no force)
        check_against_oracle(H.jiggle(dimer, seed=3, dx=0.01, dq=0.02), model.forces(fused=True)[0],
                             f"dimer {model.__name__.split('.')[-1]}")

    apart = _tiny([[-6.0, 0.0, 0.0], [6.0, 0.0, 0.0]], [q0, q0], [0, 3], np.zeros((0, 2), int), [0, 1])
    forces, _ = dna2.forces(fused=True)
    _, got = engine_eval(apart, forces)
    for t, g in got.items():
        assert np.all(g["F"] == 0.0) and np.all(g["E"] == 0.0), t


def test_run_grows_rows_mid_chunk():
    """A run whose neighbour rows outgrow their capacity: a lattice of free nucleotides
    (no bonds) imploding toward the box centre doubles its central density within the
    run. The chunk is restored, the rows grow and it runs again; the state stays finite,
    the row capacity grew, and the forces at the final state match the oracle."""
    rng = np.random.default_rng(21)
    m, a = 12, 1.8
    L = m * a
    g = (np.arange(m) - 0.5 * (m - 1)) * a
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    n = len(pos)
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1)[:, None]
    typeid = rng.integers(0, 4, size=n)
    none = -np.ones(n, dtype=int)
    system = systems.System(pos, quat, typeid, np.zeros((0, 2), dtype=int), [L] * 3, np.arange(n),
                            ["ACGT"[t] for t in typeid], none, none)
    snap = system.snapshot()
    snap.particles.velocity[:] = -0.25 * pos  # uniform compression: neighbours close in slowly
    forces, _ = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU(), seed=3)
    s.create_state_from_snapshot(snap)
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces,
                                             methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    s.run(0)
    cap0 = s.stats()["max_neigh"]
    s.run(600)
    st = s.stats()
    final = s.state.get_snapshot()
    _say(f"implosion: row capacity {cap0} -> {st['max_neigh']}, {st['nlist_builds']} builds")
    assert np.isfinite(final.particles.position).all()
    assert st["max_neigh"] > cap0
    after = systems.System(final.particles.position, final.particles.orientation, typeid, system.bonds, system.box,
                           system.strand, system.bases, none, none)
    check_against_oracle(after, dna2.forces(fused=True)[0], "after implosion")


def test_conf_trajectory_frames(tmp_path):
    """ConfTrajectory writes a frame at each multiple of its period that run() reaches;
    the last frame holds the final state, its energy line the force field's potential
    energy per nucleotide; the run split at the frames follows the unsplit run."""
    from oxdna import io
    system = H.cfg1_system(4)
    path = tmp_path / "traj.conf"
    forces, _ = dna2.forces(fused=True)
    s = _langevin_sim(system, forces, seed=3, lanes=16)
    w = sim.ConfTrajectory(str(path), 10)
    s.operations.writers.append(w)
    s.run(25)
    s.run(5)
    frames = list(io.iter_conf_frames(str(path)))
    assert [f[0] for f in frames] == [10.0, 20.0, 30.0] and w.frames == 3
    snap = s.state.get_snapshot()
    t, box, rows = frames[-1]
    L = np.asarray(snap.configuration.box[:3])
    assert np.allclose(box, L)
    part = snap.particles
    assert np.allclose(rows[:, 0:3], part.position + L / 2, atol=1e-12)
    assert np.allclose(rows[:, 3:6], io.rotate(part.orientation, [1.0, 0.0, 0.0]), atol=1e-12)
    assert np.allclose(rows[:, 6:9], io.rotate(part.orientation, [0.0, 0.0, 1.0]), atol=1e-12)
    assert np.allclose(rows[:, 9:12], part.velocity, atol=1e-12)
    assert np.allclose(rows[:, 12:15], io.angular_velocity(part.orientation, part.angmom, part.moment_inertia),
                       atol=1e-12)
    with open(path) as fh:
        lines = fh.read().splitlines()
    e_line = [x for x in lines if x.startswith("E =")][-1].split("=")[1].split()
    u = sum(s.term_energies().values()) / system.n
    assert float(e_line[1]) == pytest.approx(u, rel=1e-9, abs=1e-12)
    # the same 30 steps without writers
    forces2, _ = dna2.forces(fused=True)
    s2 = _langevin_sim(system, forces2, seed=3, lanes=16)
    s2.run(30)
    dx = np.abs(s2.state.get_snapshot().particles.position - part.position).max()
    _say("trajectory run vs plain run: max |dx|", dx)
    assert dx < 1e-6


def _rot_ke(snap):
    """Rotational kinetic energy per particle, I = 1: L_body = 1/2 (conj(q) p).v."""
    q, p = snap.particles.orientation, snap.particles.angmom
    qc = q * np.array([1.0, -1.0, -1.0, -1.0])
    L = np.stack([H.quat_mul(a, b)[1:] for a, b in zip(qc, p)]) * 0.5
    return 0.5 * np.mean(np.sum(L * L, axis=1))


def test_equipartition():
    """<KE> per particle = 3/2 kT, translational and rotational (statistical)."""
    system = systems.duplex_gas(20, 8, box=40.0, seed=21)
    forces, _ = dna2.forces(fused=True)
    kT = 0.1
    s = _langevin_sim(system, forces, seed=17, kT=kT)
    s.run(3000)
    kt, kr = [], []
    for _ in range(30):
        s.run(100)
        snap = s.state.get_snapshot()
        v = snap.particles.velocity
        kt.append(0.5 * np.mean(np.sum(v * v, axis=1)))
        kr.append(_rot_ke(snap))
    ket, ker = float(np.mean(kt)), float(np.mean(kr))
    _say(f"<KE_trans>/particle = {ket:.5f}, <KE_rot>/particle = {ker:.5f}, expected {1.5 * kT:.5f}")
    assert abs(ket - 1.5 * kT) < 0.08 * 1.5 * kT
    assert abs(ker - 1.5 * kT) < 0.08 * 1.5 * kT


def _device_nlist(system, r_list):
    """Full neighbour list in the host engine's layout from the oracle's brute-force pairs."""
    from hipmem import DeviceArray
    pairs = O.Oracle(system.pos, system.quat, system.typeid, system.box, bonds=system.bonds).pairs(r_list, False)
    n = system.n
    rows = [[] for _ in range(n)]
    for i, j in pairs:
        rows[i].append(j)
        rows[j].append(i)
    width = max(1, max(len(r) for r in rows))
    nl = np.zeros((n, width), dtype=np.uint32)
    nn = np.zeros(n, dtype=np.uint32)
    for i, r in enumerate(rows):
        nl[i, :len(r)] = sorted(r)
        nn[i] = len(r)
    head = np.arange(n, dtype=np.uint64) * width
    return DeviceArray(nn), DeviceArray(nl), DeviceArray(head)


def test_stateless_drivers_match_oracle():
    """oxdna_pair_forces / oxdna_bonded_forces on caller-owned device buffers in the
    host engine's layouts (positions with the type id in .w, (w,x,y,z) quaternions,
    n_neigh / nlist / head_list, float4 force and torque, pitched virial)."""
    from hipmem import DeviceArray
    from oxdna import _abi, _engine
    lib = _engine.load()
    system = H.jiggle(H.cfg1_system(9), seed=9)
    n = system.n
    pos4 = np.zeros((n, 4))
    pos4[:, :3] = system.pos
    pos4[:, 3] = np.frombuffer(np.asarray(system.typeid, dtype=np.int64).tobytes(), dtype=np.float64)
    d_pos, d_q, d_c = DeviceArray(pos4), DeviceArray(system.quat), DeviceArray(system.charge)
    box = _abi.Box()
    box.L[0], box.L[1], box.L[2] = system.box
    forces, nl = dna2.forces(fused=True)
    orc = O.from_system(system, forces)
    for f in forces[1:]:
        table, rcut = f._pack(["A", "G", "C", "T"])
        rcutsq = (C.c_double * 16)(*[r * r for r in rcut])
        nn, nlist, head = _device_nlist(system, max(rcut) + 0.4)
        out_f = DeviceArray(shape=(n, 4), dtype=np.float32)
        out_t = DeviceArray(shape=(n, 4), dtype=np.float32)
        out_v = DeviceArray(shape=(6, n), dtype=np.float32)
        out_e = DeviceArray(shape=(8,), dtype=np.float64)
        rc = lib.oxdna_pair_forces(f._term, out_f.ptr, out_t.ptr, out_v.ptr, n, n, d_pos.ptr, d_q.ptr, d_c.ptr,
                                   C.byref(box), nn.ptr, nlist.ptr, head.ptr, 4, C.cast(table, C.c_void_p),
                                   rcutsq, out_e.ptr, None)
        assert rc == 0, lib.oxdna_last_error(None)
        ref = orc.compute(mask=1 << f._term)
        F, T, E = out_f.get().astype(float), out_t.get().astype(float), out_e.get()
        _say(f"stateless {type(f).__name__}: dE={np.abs(E - ref['E']).max():.1e} "
             f"dF={np.abs(F[:, :3] - ref['F']).max():.1e} dT={np.abs(T[:, :3] - ref['T']).max():.1e}")
        assert np.all(np.abs(E - ref["E"]) <= E_REL * np.abs(ref["E"]) + E_PER_NT * n)
        assert np.abs(F[:, :3] - ref["F"]).max() <= F_REL * np.abs(ref["F"]).max() + F_ABS
        assert np.abs(T[:, :3] - ref["T"]).max() <= F_REL * np.abs(ref["T"]).max() + F_ABS
        assert np.abs(F[:, 3] - ref["Eh"]).max() <= F_REL * max(1.0, np.abs(ref["Eh"]).max()) + F_ABS
    # bonded, with the host engine's per-particle bond table (partner, type) + role
    ptab = forces[0]._pack(["backbone"])
    pitch = n
    blist = np.zeros((2, pitch, 2), dtype=np.uint32)
    bpos = np.zeros((2, pitch), dtype=np.uint32)
    nb = np.zeros(n, dtype=np.uint32)
    for a, b in system.bonds:
        for me, other, role in ((a, b, 0), (b, a, 1)):
            k = nb[me]
            blist[k, me] = (other, 0)
            bpos[k, me] = role
            nb[me] += 1
    d_bl, d_bp, d_nb = DeviceArray(blist), DeviceArray(bpos), DeviceArray(nb)
    out_f = DeviceArray(shape=(n, 4), dtype=np.float32)
    out_t = DeviceArray(shape=(n, 4), dtype=np.float32)
    out_v = DeviceArray(shape=(6, n), dtype=np.float32)
    split = DeviceArray(shape=(n, 4), dtype=np.float64)
    rc = lib.oxdna_bonded_forces(out_f.ptr, out_t.ptr, out_v.ptr, n, split.ptr, n, d_pos.ptr, d_q.ptr,
                                 C.byref(box), d_bl.ptr, d_bp.ptr, pitch, d_nb.ptr, ptab, 1, None)
    assert rc == 0, lib.oxdna_last_error(None)
    ref = orc.compute(mask=1 << 6)
    F, T, S = out_f.get().astype(float), out_t.get().astype(float), split.get()
    _say(f"stateless bonded: dF={np.abs(F[:, :3] - ref['F']).max():.1e} dT={np.abs(T[:, :3] - ref['T']).max():.1e} "
         f"split sums {S[:, :3].sum(axis=0)} vs {ref['E'][:3]}")
    assert np.abs(F[:, :3] - ref["F"]).max() <= F_REL * np.abs(ref["F"]).max() + F_ABS
    assert np.abs(T[:, :3] - ref["T"]).max() <= F_REL * np.abs(ref["T"]).max() + F_ABS
    assert np.allclose(S[:, :3].sum(axis=0), ref["E"][:3], rtol=E_REL, atol=E_PER_NT * n)


@pytest.mark.parametrize("model", [dna2, dna1])
def test_large_system_524k_nt(model):
    """4096 ideal 64-bp duplexes (524,288 nt) in a box of 640: the FP32 engine against
    the FP64 oracle over the whole system (per-term energies, every force and torque)."""
    import time
    t0 = time.time()
    system = systems.duplex_gas(4096, 64, box=640.0, seed=3)
    forces, nl = model.forces(fused=True)
    nl.buffer = 0.1 * max(float(f.r_cut[("A", "A")]) for f in forces[1:])  # 10 % skin
    _say(f"524k system built in {time.time() - t0:.1f} s; skin {nl.buffer:.3f}")
    s = sim.Simulation(device=sim.GPU())
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces)
    s.run(0)
    e, f4, t3, _ = s._eval(s._mask)
    ref = O.from_system(system, forces).compute()
    n = system.n
    rel = np.abs(e - ref["E"]) / np.maximum(np.abs(ref["E"]), 1e-300)
    fmax, tmax = np.abs(ref["F"]).max(), np.abs(ref["T"]).max()
    dF = np.abs(f4[:, :3] - ref["F"])
    dT = np.abs(t3 - ref["T"])
    frms = np.sqrt(np.mean(dF ** 2)) / np.sqrt(np.mean(ref["F"] ** 2))
    _say(f"[524k {model.__name__.split('.')[-1]}] E/nt={np.round(ref['E'] / n, 6).tolist()} "
         f"rel dE per term={np.round(rel, 9).tolist()} max dF/max|F|={dF.max() / fmax:.1e} "
         f"rms dF/rms F={frms:.1e} max dT/max|T|={dT.max() / tmax:.1e}")
    assert np.all(np.abs(e - ref["E"]) <= E_REL * np.abs(ref["E"]) + E_PER_NT * n)
    assert dF.max() <= F_REL * fmax + F_ABS
    assert dT.max() <= F_REL * tmax + F_ABS


@pytest.mark.parametrize("model", [dna2, rna2])
def test_thermalised_duplex_matches_oracle(model):
    """The reference's CPU-runnable case (one 20-bp duplex, box 20, Langevin dt = 0.003):
    after 1000 steps on the GPU, every term of the thermalised state against the
    oracle (energies, forces, torques, per-particle energies, virials)."""
    system = H.cfg1_system(0)
    forces, _ = model.forces()
    s = _langevin_sim(system, forces, seed=3)
    s.run(1000)
    snap = s.state.get_snapshot()
    hot = systems.System(snap.particles.position, snap.particles.orientation, system.typeid, system.bonds,
                         system.box, system.strand, system.bases, system.n3, system.n5, system.charge)
    check_against_oracle(hot, model.forces()[0], f"cfg1 after 1000 steps {model.__name__.split('.')[-1]}")
