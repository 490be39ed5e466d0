"""The host-engine surface the reference's workflows use, on the HIP engine.

The reference attaches its forces to a HOOMD-blue simulation (reference
pytest/_helpers.py:51-82, bench/benchmark.py:55-75):

    sim = Simulation(device=device.GPU(), seed=1)
    sim.create_state_from_snapshot(snap)
    forces, nlist = dna2.forces(fused=True)
    sim.operations.integrator = Integrator(dt=0.003, integrate_rotational_dof=True,
                                           forces=forces, methods=[methods.Langevin(kT=kT)])
    sim.run(steps)          # run(0) = one force evaluation (single point)
    forces[1].energy, forces[0].fene_energy, forces[2].forces, ...

This module provides that surface over one engine context (C ABI): state, bonds and
parameters are uploaded at the first run; every step evaluates all attached forces
in one fused launch; per-force read-outs (energy, forces, torques) are evaluated on
read, like the reference's energy properties. There is no CPU device: the engine is
the GPU path.
"""

import ctypes as C
import types as _types

import numpy as np

from . import _abi
from . import io
from . import _engine
from .forces import AnisotropicPair, OxDNABonded

_SLOTS = {
    _abi.TERM_BONDED: (_abi.E_FENE, _abi.E_BEXC, _abi.E_STACK),
    _abi.TERM_EXCVOL: (_abi.E_NEXC,),
    _abi.TERM_HBOND: (_abi.E_HB,),
    _abi.TERM_CROSS: (_abi.E_CRSTACK,),
    _abi.TERM_HBONDCROSS: (_abi.E_HB, _abi.E_CRSTACK),
    _abi.TERM_COAX: (_abi.E_CXSTACK,),
    _abi.TERM_DEBYE: (_abi.E_DEBYE,),
}


class _Device:
    pass


class GPU(_Device):
    """A HIP device (the only kind: the engine has no CPU path)."""

    def __init__(self, gpu_id=0):
        self.gpu_id = int(gpu_id)


class CPU(_Device):
    def __init__(self, *a, **k):
        raise RuntimeError("the oxDNA HIP engine has no CPU device; use oxdna.sim.GPU()")


device = _types.SimpleNamespace(GPU=GPU, CPU=CPU)


class Cell:
    """Cell neighbour list with skin ``buffer``; ``exclusions`` must include "bond"
    (oxDNA never pairs bonded neighbours in a nonbonded term)."""

    def __init__(self, buffer=0.4, exclusions=("bond",)):
        if "bond" not in tuple(exclusions):
            raise ValueError("oxDNA nonbonded terms require exclusions=('bond',)")
        self.buffer = float(buffer)
        self.exclusions = tuple(exclusions)


nlist = _types.SimpleNamespace(Cell=Cell)


class Langevin:
    """Langevin thermostat on translational and rotational degrees of freedom."""

    def __init__(self, filter=None, kT=0.1, default_gamma=1.0, default_gamma_r=(1.0, 1.0, 1.0)):
        self.filter = filter
        self.kT = float(kT)
        self.default_gamma = float(default_gamma)
        g = tuple(default_gamma_r) if np.iterable(default_gamma_r) else (float(default_gamma_r),) * 3
        if len(set(g)) != 1:
            raise ValueError("anisotropic rotational drag is not supported; give equal gamma_r")
        self.default_gamma_r = tuple(float(x) for x in g)


methods = _types.SimpleNamespace(Langevin=Langevin)


class Integrator:
    def __init__(self, dt, integrate_rotational_dof=True, forces=(), methods=()):
        self.dt = float(dt)
        self.integrate_rotational_dof = bool(integrate_rotational_dof)
        self.forces = list(forces)
        self.methods = list(methods)
        if len(self.methods) > 1:
            raise ValueError("one integration method (Langevin) per integrator")


class ConfTrajectory:
    """Writes an oxDNA configuration frame to `filename` at every timestep that run()
    reaches with timestep % period == 0: the line `t = ...`, the box, `E = Etot U K`
    (per nucleotide: potential energy of the attached force field, kinetic energy of
    translation and rotation), then per particle the position (shifted into [0, L]),
    a1, a3, velocity and lab-frame angular velocity. mode "w" starts the file at the
    first frame, "a" appends to it. Frames read back with io.iter_conf_frames."""

    def __init__(self, filename, period, mode="w"):
        if int(period) < 1:
            raise ValueError("period must be a positive number of steps")
        if mode not in ("w", "a"):
            raise ValueError('mode must be "w" or "a"')
        self.filename = filename
        self.period = int(period)
        self._append = mode == "a"
        self.frames = 0

    def write(self, simulation):
        snap = simulation.state.get_snapshot()
        part = snap.particles
        n = part.N
        pot = float(sum(simulation.term_energies().values()))
        kin = io.kinetic_energy(part.mass, part.velocity, part.orientation, part.angmom, part.moment_inertia)
        omega = io.angular_velocity(part.orientation, part.angmom, part.moment_inertia)
        io.write_configuration(self.filename, simulation.timestep, snap.configuration.box[:3], part.position,
                               part.orientation, part.velocity, omega,
                               energy=((pot + kin) / n, pot / n, kin / n), append=self._append)
        self._append = True
        self.frames += 1


class _Operations:
    def __init__(self, sim):
        self._sim = sim
        self._integrator = None
        self.writers = []  # ConfTrajectory instances, called as run() reaches their periods

    @property
    def integrator(self):
        return self._integrator

    @integrator.setter
    def integrator(self, integ):
        self._integrator = integ
        self._sim._attached = False


class _State:
    def __init__(self, sim):
        self._sim = sim

    def get_snapshot(self):
        return self._sim._get_snapshot()

    @property
    def N_particles(self):
        return self._sim._snap.particles.N


class Simulation:
    """One system on one GPU, driven through the engine's C ABI."""

    def __init__(self, device=None, seed=0, use_graph=True, lanes_per_particle=0):
        """lanes_per_particle: 4, 8, 16 or 32 lanes of the force kernel per particle, or 0
        to autotune (the choice fixes the summation order: pin it for bit-reproducible
        runs across processes)."""
        self.device = device if device is not None else GPU()
        if not isinstance(self.device, GPU):
            raise RuntimeError("the oxDNA HIP engine runs on a GPU device only")
        self.seed = int(seed)
        self.use_graph = bool(use_graph)
        self._lib = _engine.load()
        self._ctx = C.c_void_p()
        _engine.check(self._lib, None, self._lib.oxdna_ctx_create(self.device.gpu_id, None,
                                                                  C.byref(self._ctx)), "ctx_create")
        if lanes_per_particle:
            self._check(self._lib.oxdna_set_lanes_per_particle(self._ctx, int(lanes_per_particle)),
                        "set_lanes_per_particle")
        self._snap = None
        self._attached = False
        self._mask = 0
        self._cache = {}
        self.timestep = 0
        self.operations = _Operations(self)
        self.state = _State(self)

    def __del__(self):
        lib = getattr(self, "_lib", None)
        ctx = getattr(self, "_ctx", None)
        if lib is not None and ctx is not None and ctx.value:
            lib.oxdna_ctx_destroy(ctx)
            self._ctx = C.c_void_p()

    def _check(self, rc, what):
        _engine.check(self._lib, self._ctx, rc, what)

    # ------------------------------------------------------------------ state
    def create_state_from_snapshot(self, snap):
        self._snap = snap.copy()
        self._upload_state()
        self._attached = False

    def _upload_state(self):
        p = self._snap.particles
        n = p.N
        pos = np.ascontiguousarray(p.position, dtype=np.float64)
        quat = np.ascontiguousarray(p.orientation, dtype=np.float64)
        typ = np.ascontiguousarray(p.typeid, dtype=np.int32)
        chg = np.ascontiguousarray(p.charge, dtype=np.float64)
        box = _abi.Box()
        box.L[0], box.L[1], box.L[2] = self._snap.configuration.box[:3]
        dp = C.POINTER(C.c_double)
        self._check(self._lib.oxdna_set_state(self._ctx, n, pos.ctypes.data_as(dp), quat.ctypes.data_as(dp),
                                              typ.ctypes.data_as(C.POINTER(C.c_int32)),
                                              chg.ctypes.data_as(dp), C.byref(box)), "set_state")
        vel = np.ascontiguousarray(p.velocity, dtype=np.float64)
        ang = np.ascontiguousarray(p.angmom, dtype=np.float64)
        self._check(self._lib.oxdna_set_momenta(self._ctx, vel.ctypes.data_as(dp), ang.ctypes.data_as(dp)),
                    "set_momenta")
        b = self._snap.bonds
        ab = np.ascontiguousarray(np.asarray(b.group, dtype=np.int32).reshape(-1, 2))
        bt = np.ascontiguousarray(np.asarray(b.typeid, dtype=np.int32))
        ip = C.POINTER(C.c_int32)
        self._check(self._lib.oxdna_set_bonds(self._ctx, int(b.N), ab.ctypes.data_as(ip), bt.ctypes.data_as(ip)),
                    "set_bonds")
        self._cache.clear()

    def _get_snapshot(self):
        n = self._snap.particles.N
        pos, quat = np.empty((n, 3)), np.empty((n, 4))
        vel, ang = np.empty((n, 3)), np.empty((n, 4))
        dp = C.POINTER(C.c_double)
        self._check(self._lib.oxdna_get_state(self._ctx, pos.ctypes.data_as(dp), quat.ctypes.data_as(dp),
                                              vel.ctypes.data_as(dp), ang.ctypes.data_as(dp)), "get_state")
        s = self._snap.copy()
        s.particles.position[:] = pos
        s.particles.orientation[:] = quat
        s.particles.velocity[:] = vel
        s.particles.angmom[:] = ang
        return s

    # ----------------------------------------------------------------- attach
    def _attach(self):
        integ = self.operations.integrator
        if integ is None:
            raise RuntimeError("set sim.operations.integrator before running")
        if self._snap is None:
            raise RuntimeError("create the state before running")
        types = list(self._snap.particles.types)
        seen = {}
        mask = 0
        buffer = 0.0
        for f in integ.forces:
            if f._term in seen:
                raise ValueError(f"two forces of kind {type(f).__name__} attached; the engine "
                                 "evaluates one parameter set per term")
            seen[f._term] = f
            mask |= 1 << f._term
            if isinstance(f, OxDNABonded):
                table = f._pack(list(self._snap.bonds.types))
                self._check(self._lib.oxdna_set_bond_params(self._ctx, len(table), table), "set_bond_params")
            elif isinstance(f, AnisotropicPair):
                table, rcut = f._pack(types)
                self._check(self._lib.oxdna_set_pair_params(self._ctx, f._term, len(types),
                                                            C.cast(table, C.c_void_p), rcut),
                            "set_pair_params")
                buffer = max(buffer, f.nlist.buffer)
            else:
                raise TypeError(f"unsupported force {f!r}")
            f._simulation = self
        for t in range(_abi.N_TERMS):
            if t != _abi.TERM_BONDED and t not in seen:
                self._lib.oxdna_set_pair_params(self._ctx, t, len(types), None, None)
        self._check(self._lib.oxdna_set_nlist(self._ctx, buffer), "set_nlist")
        lang = [m for m in integ.methods if isinstance(m, Langevin)]
        if lang:
            m = lang[0]
            self._check(self._lib.oxdna_set_langevin(self._ctx, integ.dt, m.kT, m.default_gamma,
                                                     m.default_gamma_r[0], self.seed), "set_langevin")
        self._mask = mask
        self._attached = True
        self._cache.clear()

    # -------------------------------------------------------------------- run
    def run(self, steps):
        if not self._attached:
            self._attach()
        steps = int(steps)
        if steps == 0:
            self._check(self._lib.oxdna_compute(self._ctx, self._mask, 1), "compute")
        else:
            if not self.operations.integrator.methods:
                raise RuntimeError("running steps needs an integration method (methods.Langevin)")
            left = steps
            while left > 0:
                # run up to the next timestep at which a writer is due
                chunk = left
                for w in self.operations.writers:
                    chunk = min(chunk, w.period - self.timestep % w.period)
                self._check(self._lib.oxdna_run(self._ctx, chunk, self._mask, 1 if self.use_graph else 0), "run")
                self.timestep += chunk
                left -= chunk
                self._cache.clear()
                for w in self.operations.writers:
                    if self.timestep % w.period == 0:
                        w.write(self)
        self._cache.clear()

    def synchronize(self):
        self._check(self._lib.oxdna_synchronize(self._ctx), "synchronize")

    def tune_nlist_buffer(self, buffers=None, steps=2000):
        """Neighbour-list skin autotuner (the reference fixes buffer 0.4,
        oxdna/model/dna2.py:207-208). From the current state, runs `steps` Langevin
        steps at each candidate buffer (default 0.1, 0.2, ..., 0.8), keeps the one with
        the most steps/s, sets it on the attached neighbour lists and restores the state
        and timestep. Returns {buffer: steps/s}. Tune on an equilibrated state: the
        rebuild rate it measures is that state's."""
        if not self._attached:
            self._attach()
        if not self.operations.integrator.methods:
            raise RuntimeError("tuning the skin runs steps: attach an integration method (methods.Langevin)")
        cand = None if buffers is None else np.ascontiguousarray([float(b) for b in buffers], dtype=np.float64)
        n = 8 if cand is None else len(cand)
        rates = np.zeros(n)
        chosen = C.c_double()
        dp = C.POINTER(C.c_double)
        self._check(self._lib.oxdna_tune_nlist_buffer(
            self._ctx, self._mask, None if cand is None else cand.ctypes.data_as(dp), n, int(steps),
            rates.ctypes.data_as(dp), C.byref(chosen)), "tune_nlist_buffer")
        for f in self.operations.integrator.forces:
            if isinstance(f, AnisotropicPair):
                f.nlist.buffer = chosen.value
        self._cache.clear()
        grid = [0.1 * k for k in range(1, 9)] if cand is None else list(cand)
        return {round(float(b), 6): float(r) for b, r in zip(grid, rates)}

    def tuning(self):
        """The autotuner's current launch parameters and skin."""
        t = _abi.Tuning()
        self._check(self._lib.oxdna_get_tuning_ex(self._ctx, C.byref(t)), "get_tuning_ex")
        return {name: getattr(t, name) for name, _ in t._fields_}

    def stats(self):
        nb, mx, st = C.c_uint64(), C.c_uint32(), C.c_uint64()
        self._check(self._lib.oxdna_get_stats(self._ctx, C.byref(nb), C.byref(mx), C.byref(st)), "get_stats")
        return dict(nlist_builds=nb.value, max_neigh=mx.value, step=st.value)

    # --------------------------------------------------------------- read-out
    def _eval(self, term_mask):
        key = ("e", term_mask)
        if key not in self._cache:
            self._check(self._lib.oxdna_compute(self._ctx, term_mask, 1), "compute")
            e = np.zeros(8)
            self._check(self._lib.oxdna_get_term_energies(self._ctx, e.ctypes.data_as(C.POINTER(C.c_double))),
                        "get_term_energies")
            n = self._snap.particles.N
            f4, t3, v6 = np.empty((n, 4)), np.empty((n, 3)), np.empty((n, 6))
            dp = C.POINTER(C.c_double)
            self._check(self._lib.oxdna_get_forces(self._ctx, f4.ctypes.data_as(dp), t3.ctypes.data_as(dp),
                                                   v6.ctypes.data_as(dp)), "get_forces")
            self._cache[key] = (e, f4, t3, v6)
        return self._cache[key]

    def _term_energies(self, force):
        return self._eval(1 << force._term)[0]

    def _force_energy(self, force):
        e = self._term_energies(force)
        return float(sum(e[s] for s in _SLOTS[force._term]))

    def _force_arrays(self, force):
        e, f4, t3, v6 = self._eval(1 << force._term)
        return f4[:, :3].copy(), t3.copy(), f4[:, 3].copy(), v6.copy()

    def term_energies(self):
        """All per-term energies of the attached force field (split_energy order)."""
        if not self._attached:
            self._attach()
        return dict(zip(_abi.ENERGY_NAMES, self._eval(self._mask)[0]))

    def neighbor_pairs(self):
        """The engine's current neighbour list as sorted (i, j) pairs, i < j."""
        n = self._snap.particles.N
        cap = max(1, n * 64)
        while True:
            buf = np.empty((cap, 2), dtype=np.int32)
            cnt = C.c_size_t()
            rc = self._lib.oxdna_get_nlist_pairs(self._ctx, buf.ctypes.data_as(C.POINTER(C.c_int32)), cap,
                                                 C.byref(cnt))
            if rc == 0:
                return buf[:cnt.value].copy()
            if cnt.value > cap:
                cap = cnt.value
                continue
            self._check(rc, "get_nlist_pairs")
