"""GPU parity at the workloads bench.py times, against the FP64 oracle.

* The 8,000-nt oxDNA2 system (200 ideal 20-bp duplexes, box 200, fused HB + CS) at 16
  lanes per particle: 500 force workgroups, the grid shape of the timed run step (its
  second dispatch batch raises its wave priority), checked at its initial state and
  after a 1,000-step graph run:
    - energies, forces, torques, per-particle energies and virials of a single-point
      evaluation (the energy + virial instantiation);
    - forces and torques of the forces-only FAR2 launch that ends every run;
    - forces and torques of the FUSED run-step launch itself (forces + Langevin
      integrator in one kernel, which stores neither): with kT = gamma = gamma_r = 0
      the integrator is velocity Verlet plus the NO_SQUISH rotor, linear in the force
      and torque, so a 2-step run and a 1-step run from the same state give them back
      exactly (FP64 velocities and momenta; the rotor is inverted on the host).
* The 512-hairpin and 256-cloverleaf (k = 0.02 bias) ensembles, replica by replica.

Tolerances are the one-system ones: per-term energy |dE| <= 1e-4 |E_ref| + 1e-5 N,
per-particle force / torque |d| <= 2e-4 max|ref| + 2e-4.
"""

import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from oxdna import ensemble, sim, systems
from oxdna.model import dna1, dna2, rna2

from test_gpu_ensemble import _check_replicas
from test_gpu_parity import check_against_oracle

pytestmark = pytest.mark.gpu

F_REL, F_ABS = 2e-4, 2e-4
DT = 0.003
KT = dna1.kT_from_temperature(296.15)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------- host quaternion algebra
def qmul(a, b):
    """Row-wise quaternion products, (w, x, y, z) layout."""
    aw, ax, ay, az = a.T
    bw, bx, by, bz = b.T
    return np.stack([aw * bw - ax * bx - ay * by - az * bz,
                     aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw], axis=1)


def qconj(a):
    return a * np.array([1.0, -1.0, -1.0, -1.0])


def pure(v):
    return np.concatenate([np.zeros((len(v), 1)), v], axis=1)


def permute(k, q):
    w, x, y, z = q.T
    if k == 1:
        return np.stack([-x, w, z, -y], axis=1)
    if k == 2:
        return np.stack([-y, -z, w, x], axis=1)
    return np.stack([-z, y, -x, w], axis=1)


def rotor(k, tau, q, p, inertia=1.0):
    """Free rotation about body axis k for time tau (NO_SQUISH, Miller et al. 2002).
    zeta = p . P_k q / (4 I) is invariant under it, so rotor(k, -tau) undoes rotor(k, tau)."""
    pq, pp = permute(k, q), permute(k, p)
    zeta = np.sum(p * pq, axis=1) / (4.0 * inertia)
    c, s = np.cos(zeta * tau)[:, None], np.sin(zeta * tau)[:, None]
    return c * q + s * pq, c * p + s * pp


def unrotate(q, p, dt):
    """Inverse of the step's rotor sequence (axes 3, 2, 1, 2, 3 for dt/2, dt/2, dt,
    dt/2, dt/2)."""
    for k, tau in ((3, -0.5 * dt), (2, -0.5 * dt), (1, -dt), (2, -0.5 * dt), (3, -0.5 * dt)):
        q, p = rotor(k, tau, q, p)
    return q, p


# ---------------------------------------------------------------- engine helpers
def _nve_sim(system, lanes, pos=None, quat=None, vel=None, angmom=None):
    forces, _ = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU(), seed=7, lanes_per_particle=lanes)
    snap = system.snapshot()
    for name, val in (("position", pos), ("orientation", quat), ("velocity", vel), ("angmom", angmom)):
        if val is not None:
            getattr(snap.particles, name)[:] = val
    s.create_state_from_snapshot(snap)
    s.operations.integrator = sim.Integrator(
        dt=DT, forces=forces, methods=[sim.Langevin(kT=0.0, default_gamma=0.0, default_gamma_r=(0.0, 0.0, 0.0))])
    return s, forces


def _raw_forces(s):
    """Force and torque buffers as the last launch left them (no new evaluation)."""
    n = s._snap.particles.N
    f4, t3, v6 = np.empty((n, 4)), np.empty((n, 3)), np.empty((n, 6))
    dp = C.POINTER(C.c_double)
    s._check(s._lib.oxdna_get_forces(s._ctx, f4.ctypes.data_as(dp), t3.ctypes.data_as(dp), v6.ctypes.data_as(dp)),
             "get_forces")
    return f4[:, :3], t3


def _state(s):
    p = s.state.get_snapshot().particles
    return p.position.copy(), p.orientation.copy(), p.velocity.copy(), p.angmom.copy()


def _at(system, pos, quat):
    return systems.System(pos, quat, system.typeid, system.bonds, system.box, system.strand, system.bases,
                          system.n3, system.n5, system.charge)


def _cmp(label, F, T, ref):
    fmax, tmax = np.abs(ref["F"]).max(), np.abs(ref["T"]).max()
    dF, dT = np.abs(F - ref["F"]).max(), np.abs(T - ref["T"]).max()
    _say(f"  {label}: dF={dF:.2e} (max|F| {fmax:.2e}) dT={dT:.2e} (max|T| {tmax:.2e})")
    assert dF <= F_REL * fmax + F_ABS, f"{label}: force error {dF}"
    assert dT <= F_REL * tmax + F_ABS, f"{label}: torque error {dT}"


def _check_run_kernels(system, label, vel=None, angmom=None, lanes=16):
    """The forces-only FAR2 launch and the FUSED run-step launch against the oracle, from
    `system` with the given momenta (see the module docstring)."""
    hm = 0.5 * DT
    a, forces = _nve_sim(system, lanes, vel=vel, angmom=angmom)
    a.run(1)                                  # k_step1, then the last step (FAR2, + second half)
    x1, q1, v1e, p1e = _state(a)
    F1a, T1a = _raw_forces(a)                 # FAR2 at state 1
    b, _ = _nve_sim(system, lanes, vel=vel, angmom=angmom)
    b.run(2)                                  # k_step1, FUSED step 1, last step 2 (FAR2)
    x2, q2, v2e, p2e = _state(b)
    F2b, T2b = _raw_forces(b)                 # FAR2 at state 2
    tuning = C.c_int()
    b._check(b._lib.oxdna_get_tuning(b._ctx, C.byref(tuning), None), "get_tuning")
    assert tuning.value == lanes
    blocks = -(-system.n * lanes // 256)
    _say(f"[{label}] n={system.n}, {lanes} lanes per particle: {blocks} force workgroups")
    # translation: v_half = v1e - hm F1a; FUSED: v_c = v_half + 2 hm F1; v2e = v_c + hm F2b
    v_half = v1e - hm * F1a
    F1 = (v2e - hm * F2b - v_half) / (2.0 * hm)
    # rotation: p_b = p1e - dt (0, T1a) (x) q1; FUSED: p_c = p_b + 2 dt (0, T1) (x) q1,
    # rotor (q1, p_c) -> (q2, p_d); p2e = p_d + dt (0, T2b) (x) q2
    p_b = p1e - DT * qmul(pure(T1a), q1)
    p_d = p2e - DT * qmul(pure(T2b), q2)
    q1r, p_c = unrotate(q2, p_d, DT)
    assert np.abs(q1r - q1).max() < 1e-10
    T1 = qmul((p_c - p_b) / (2.0 * DT), qconj(q1))
    assert np.abs(T1[:, 0]).max() < 1e-6 * max(1.0, np.abs(T1).max())
    T1 = T1[:, 1:]
    ref1 = O.from_system(_at(system, x1, q1), forces).compute()
    ref2 = O.from_system(_at(system, x2, q2), forces).compute()
    _cmp("FUSED run step, state 1", F1, T1, ref1)
    _cmp("forces-only FAR2, state 1", F1a, T1a, ref1)
    _cmp("forces-only FAR2, state 2", F2b, T2b, ref2)


def _cfg2():
    return systems.duplex_gas(200, 20, box=200.0, seed=1)


def test_bench_8k_workload_initial_state():
    system = _cfg2()
    forces, _ = dna2.forces(fused=True)
    check_against_oracle(system, forces, "8k nt initial, single point", lanes=16)
    _check_run_kernels(system, "8k nt initial")


def test_bench_8k_workload_after_1000_graph_steps():
    system = _cfg2()
    forces, _ = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU(), seed=1, lanes_per_particle=16)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=DT, forces=forces, methods=[sim.Langevin(kT=KT)])
    s.run(1000)
    pos, quat, vel, ang = _state(s)
    st = s.stats()
    _say(f"after 1000 graph steps: {st['nlist_builds']} neighbour-list builds, max row {st['max_neigh']}")
    assert st["nlist_builds"] > 1
    hot = _at(system, pos, quat)
    # the forces-only launch that ended the run, at the run's final state
    F, T = _raw_forces(s)
    ref = O.from_system(hot, forces).compute()
    _cmp("forces-only FAR2 that ended the 1000-step run", F, T, ref)
    check_against_oracle(hot, forces, "8k nt after 1000 steps, single point", lanes=16)
    _check_run_kernels(hot, "8k nt after 1000 steps", vel=vel, angmom=ang)


@pytest.mark.parametrize("lanes", [4, 8, 32])
def test_fused_run_step_other_layouts(lanes):
    """The FUSED run-step launch at the other particles-per-workgroup layouts (64, 32 and
    8 particles per workgroup; its epilogue integrates a workgroup's particles on the
    first wavefront, one lane each) against the oracle: a 2,400-nt gas of 10-bp duplexes
    (rows of at most ~21 entries, which 64 particles per workgroup hold in LDS; longer
    rows make the engine take 32 lanes instead) after 300 thermal steps, with those
    steps' momenta."""
    system = systems.duplex_gas(120, 10, box=80.0, seed=11)
    forces, _ = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU(), seed=2, lanes_per_particle=lanes)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=DT, forces=forces, methods=[sim.Langevin(kT=KT)])
    s.run(300)
    pos, quat, vel, ang = _state(s)
    _check_run_kernels(_at(system, pos, quat), f"2,400-nt gas after 300 steps, {lanes} lanes", vel=vel,
                       angmom=ang, lanes=lanes)


def test_bench_512_hairpins():
    reps = systems.hairpin_ensemble(512, seed=3, box=30.0)
    forces, _ = dna2.forces(fused=True)
    _check_replicas(reps, forces, "512 hairpins (bench ensemble)")


def test_bench_256_cloverleaves_with_bias():
    reps, ends, k, d0 = systems.cloverleaf_ensemble(256, seed=5, box=64.0, d_range=(2.0, 30.0), k_bias=0.02)
    forces, _ = rna2.forces(fused=True)
    _check_replicas(reps, forces, "256 cloverleaves + bias k=0.02 (bench ensemble)",
                    bias=ensemble.HarmonicBias(ends, k, d0))
