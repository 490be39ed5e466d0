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


def engine_eval(system, forces):
    """Attach forces to a Simulation, run(0), read per-force outputs."""
    s = sim.Simulation(device=sim.GPU())
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces)
    s.run(0)
    out = {}
    for f in forces:
        e = s._term_energies(f)
        F, T, Eh, V = f.forces, f.torques, f.energies, f.virials
        out[f._term] = dict(E=e, F=F, T=T, Eh=Eh, V=V)
    return s, out


def check_against_oracle(system, forces, label):
    orc = O.from_system(system, forces)
    _, got = engine_eval(system, forces)
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
        _say(f"[{label}] {name}: E_ref/nt={ref['E'] / n} dE={dE.max():.2e} dF={dF:.2e}/{fmax:.2e} "
             f"dT={dT:.2e}/{tmax:.2e} dEh={dEh:.2e} dV={dV:.2e}/{vmax:.2e}")
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


def _langevin_sim(system, forces, seed=5, use_graph=True, kT=None):
    s = sim.Simulation(device=sim.GPU(), seed=seed, use_graph=use_graph)
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
    system = H.cfg1_system(4)
    out = []
    for g in (True, False):
        forces, _ = dna2.forces(fused=True)
        s = _langevin_sim(system, forces, seed=9, use_graph=g)
        s.run(100)
        out.append(s.state.get_snapshot().particles.position.copy())
    _say("graph vs direct max |dx|", np.abs(out[0] - out[1]).max())
    assert np.array_equal(out[0], out[1])


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
