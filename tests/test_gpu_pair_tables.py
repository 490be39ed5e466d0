"""Per-type-pair parameter tables on the GPU against the FP64 oracle.

The reference's pair evaluators take one parameter struct per type pair (HOOMD's
param_type per (ti, tj); every nonbonded force keys its params by pair, reference
oxdna/forces.py len_keys=2). Here the tables differ between type pairs in more than the
hydrogen-bonding epsilon: excluded-volume epsilon and base-base sigma, cross-stacking
(and fused HB + cross-stacking) spring constant, coaxial-stacking spring constant and
Debye-Hueckel prefactor, each on some type pairs only. The engine must pick each pair's
own set on the context path (fused and unfused), in the replica kernel and in the
stateless driver. Tolerances as in test_gpu_parity.
"""

import ctypes as C

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import ensemble, sim, systems
from oxdna.model import dna2

from test_gpu_ensemble import _check_replicas
from test_gpu_parity import E_PER_NT, E_REL, F_ABS, F_REL, _device_nlist, _say, check_against_oracle

pytestmark = pytest.mark.gpu


def _scaled(d, key, k, factor):
    v = list(d[key])
    v[k] *= factor
    d[key] = tuple(v)


def hetero_forces(fused):
    forces, nl = dna2.forces(fused=fused)
    by = {type(f).__name__: f for f in forces}
    ev = by["ExcludedVolume"]
    v = dict(ev.params[("A", "A")])
    v["eps"] *= 1.6
    ev.params[("A", "T")] = v
    v = dict(ev.params[("A", "A")])
    _scaled(v, "base_base", 0, 1.08)  # sigma of the base-base site pair
    ev.params[("G", "C")] = v
    for name in ("CrossStacking", "HBondCross"):
        if name in by:
            f = by[name]
            v = dict(f.params[("A", "A")])
            _scaled(v, "f2", 0, 1.7)       # spring constant K
            f.params[("A", "G")] = v
            v = dict(f.params[("A", "A")])
            _scaled(v, "f2", 0, 0.6)
            f.params[("C", "T")] = v
    cx = by["CoaxialStacking"]
    v = dict(cx.params[("A", "A")])
    _scaled(v, "f2", 0, 1.5)
    for key in (("A", "A"), ("A", "G"), ("A", "C"), ("A", "T"), ("G", "G"), ("G", "C"), ("G", "T")):
        cx.params[key] = v
    dh = by["DebyeHuckel"]
    v = dict(dh.params[("A", "A")])
    _scaled(v, "debye", 1, 2.0)            # prefactor
    dh.params[("C", "C")] = v
    return forces


def _system():
    return H.jiggle(systems.duplex_gas(6, 20, box=30.0, seed=11, min_dist=1.2), seed=3, dx=0.03, dq=0.05)


def _nicked():
    import census
    return H.jiggle(census.nicked_helix(n_bp=8, copies=3), seed=8, dx=0.02, dq=0.03)


@pytest.mark.parametrize("fused", [True, False])
def test_per_type_pair_tables_context(fused):
    changed = np.zeros(8, dtype=bool)
    for label, system in (("duplex gas", _system()), ("nicked helix", _nicked())):
        forces = hetero_forces(fused)
        base, _ = dna2.forces(fused=fused)
        e_mod = O.from_system(system, forces).compute()["E"]
        e_base = O.from_system(system, base).compute()["E"]
        _say(f"{label} fused={fused}: per-term energy change from the per-pair sets "
             f"{np.round(e_mod - e_base, 5).tolist()}")
        changed |= np.abs(e_mod - e_base) > 1e-6
        check_against_oracle(system, forces, f"per-pair tables, {label}, fused={fused}")
    # the modified tables reach excluded volume, cross-stacking, coaxial stacking and
    # Debye-Hueckel (the test is not vacuous)
    assert changed[[3, 5, 6, 7]].all(), changed


def test_per_type_pair_tables_run_and_replicas():
    system = _system()
    forces = hetero_forces(True)
    # a short run on the fused run-step path, then the state against the oracle
    s = sim.Simulation(device=sim.GPU(), seed=2)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(dt=0.003, forces=forces, methods=[sim.Langevin(kT=0.0987)])
    s.run(300)
    snap = s.state.get_snapshot()
    hot = systems.System(snap.particles.position, snap.particles.orientation, system.typeid, system.bonds,
                         system.box, system.strand, system.bases, system.n3, system.n5, system.charge)
    check_against_oracle(hot, hetero_forces(True), "per-pair tables after 300 steps")
    reps = [H.jiggle(h, seed=30 + k, dx=0.03, dq=0.05) for k, h in enumerate(systems.hairpin_ensemble(8, seed=4))]
    _check_replicas(reps, hetero_forces(True), "per-pair tables, replicas")


def test_per_type_pair_tables_stateless_driver():
    from hipmem import DeviceArray
    from oxdna import _abi, _engine
    lib = _engine.load()
    system = _system()
    n = system.n
    pos4 = np.zeros((n, 4))
    pos4[:, :3] = system.pos
    pos4[:, 3] = np.frombuffer(np.asarray(system.typeid, dtype=np.int64).tobytes(), dtype=np.float64)
    d_pos, d_q, d_c = DeviceArray(pos4), DeviceArray(system.quat), DeviceArray(system.charge)
    box = _abi.Box()
    box.L[0], box.L[1], box.L[2] = system.box
    forces = hetero_forces(True)
    orc = O.from_system(system, forces)
    for f in forces[1:]:
        table, rcut = f._pack(["A", "G", "C", "T"])
        rcutsq = (C.c_double * 16)(*[r * r for r in rcut])
        nn, nlist, head = _device_nlist(system, max(rcut) + 0.4)
        out_f = DeviceArray(shape=(n, 4), dtype=np.float32)
        out_t = DeviceArray(shape=(n, 4), dtype=np.float32)
        out_e = DeviceArray(shape=(8,), dtype=np.float64)
        for entry in ("oxdna_pair_forces", "oxdna_pair_forces_noalloc"):
            rc = getattr(lib, entry)(f._term, out_f.ptr, out_t.ptr, None, 0, n, d_pos.ptr, d_q.ptr, d_c.ptr,
                                     C.byref(box), nn.ptr, nlist.ptr, head.ptr, 4, C.cast(table, C.c_void_p), rcutsq,
                                     out_e.ptr, None)
            assert rc == 0, lib.oxdna_last_error(None)
            ref = orc.compute(mask=1 << f._term)
            F, T, E = out_f.get().astype(float), out_t.get().astype(float), out_e.get()
            _say(f"{entry} {type(f).__name__}: dE={np.abs(E - ref['E']).max():.1e} "
                 f"dF={np.abs(F[:, :3] - ref['F']).max():.1e}")
            assert np.all(np.abs(E - ref["E"]) <= E_REL * np.abs(ref["E"]) + E_PER_NT * n)
            assert np.abs(F[:, :3] - ref["F"]).max() <= F_REL * np.abs(ref["F"]).max() + F_ABS
            assert np.abs(T[:, :3] - ref["T"]).max() <= F_REL * np.abs(ref["T"]).max() + F_ABS
