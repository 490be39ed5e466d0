"""Batched replica ensembles: many independent small systems on one GPU.

The reference runs one system per host-engine simulation; hairpin ensembles,
duplex libraries and umbrella windows are then many separate runs. Here they are
one ``ReplicaEnsemble``: the replicas' particles are concatenated in one engine
context, each replica keeps its own periodic box, no pair or bond crosses replicas,
and every chunk of Langevin steps is one persistent launch with one workgroup per
replica (include/oxdna_hip.h, oxdna_set_replicas).

    ens = ReplicaEnsemble(device=sim.GPU(), seed=7)
    ens.create_state_from_snapshots([s.snapshot() for s in systems])
    ens.operations.integrator = sim.Integrator(dt=0.003, forces=forces,
                                               methods=[sim.Langevin(kT=kT)])
    ens.bias = HarmonicBias(ends, k, d0)        # optional, per replica
    ens.run(10_000)
    ens.replica_energies()                      # (n_replicas, 9): 8 terms + bias

Replica r draws its Langevin noise with key ``seed + r``: a one-system
``Simulation(seed=seed + r)`` holding replica r alone follows the same random
stream.
"""

import ctypes as C

import numpy as np

from . import _abi, _engine, io, sim

REPLICA_ENERGY_NAMES = list(_abi.ENERGY_NAMES) + ["bias"]
MAX_REPLICA_PARTICLES = 512


class HarmonicBias:
    """Per-replica harmonic bias k (|r_e - r_s| - d0)^2 on the centres of two
    particles of each replica (``ends`` = (s, e) local indices). Not a term of the
    reference: a new force for umbrella-style sampling of, e.g., end-to-end distance."""

    def __init__(self, ends, k, d0):
        self.ends = np.ascontiguousarray(np.asarray(ends, dtype=np.int32).reshape(-1, 2))
        n = len(self.ends)
        self.k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=float), (n,)).copy())
        self.d0 = np.ascontiguousarray(np.broadcast_to(np.asarray(d0, dtype=float), (n,)).copy())

    def energy(self, positions, boxes, offsets):
        """The bias energies for host positions (FP64), replica by replica."""
        out = np.zeros(len(self.ends))
        for r, (s, e) in enumerate(self.ends):
            o = offsets[r]
            d = positions[o + e] - positions[o + s]
            L = np.asarray(boxes[r])
            d = d - L * np.rint(d / L)
            out[r] = self.k[r] * (np.linalg.norm(d) - self.d0[r]) ** 2
        return out


def concatenate_snapshots(snaps):
    """One snapshot holding every replica in order, plus the particle offsets and the
    per-replica boxes. Particle and bond types must agree across the replicas."""
    if not snaps:
        raise ValueError("no replicas")
    types = list(snaps[0].particles.types)
    btypes = list(snaps[0].bonds.types)
    offsets = [0]
    boxes = []
    for s in snaps:
        if list(s.particles.types) != types or list(s.bonds.types) != btypes:
            raise ValueError("every replica must use the same particle and bond types")
        n = s.particles.N
        if n > MAX_REPLICA_PARTICLES:
            raise ValueError(f"a replica has {n} particles; at most {MAX_REPLICA_PARTICLES} per replica")
        offsets.append(offsets[-1] + n)
        boxes.append(list(s.configuration.box[:3]))
    out = io.Snapshot()
    out.particles.resize(offsets[-1])
    out.particles.types = types
    for name in ("position", "typeid", "mass", "moment_inertia", "orientation", "charge", "velocity", "angmom"):
        getattr(out.particles, name)[:] = np.concatenate([getattr(s.particles, name) for s in snaps])
    groups = [np.asarray(s.bonds.group, dtype=np.int32).reshape(-1, 2) + o for s, o in zip(snaps, offsets)]
    out.bonds.types = btypes
    out.bonds.group = np.concatenate(groups).astype(np.int32)
    out.bonds.typeid = np.concatenate([np.asarray(s.bonds.typeid, dtype=np.int32) for s in snaps])
    out.bonds.N = len(out.bonds.group)
    big = np.max(np.asarray(boxes), axis=0)
    out.configuration.box = [big[0], big[1], big[2], 0.0, 0.0, 0.0]
    return out, np.asarray(offsets, dtype=np.uint32), np.asarray(boxes, dtype=float)


class ReplicaEnsemble(sim.Simulation):
    """Independent replicas of possibly different sizes, run together."""

    def __init__(self, device=None, seed=0):
        super().__init__(device=device, seed=seed, use_graph=False)
        self._offsets = None
        self._boxes = None
        self._bias = None

    @property
    def n_replicas(self):
        return 0 if self._offsets is None else len(self._offsets) - 1

    @property
    def offsets(self):
        return self._offsets.copy()

    @property
    def boxes(self):
        return self._boxes.copy()

    def create_state_from_snapshots(self, snaps):
        snap, self._offsets, self._boxes = concatenate_snapshots(list(snaps))
        self.create_state_from_snapshot(snap)

    def set_tiling(self, workgroup=0, depth_divisor=1):
        """Pins the replica kernel's shape instead of autotuning it: `workgroup` threads
        per replica (64, 128, 256 or 512, at least the largest replica) and pair records
        per pass of the LDS share's capacity / `depth_divisor` (1, 2 or 4). workgroup=0
        hands both back to the autotuner. Every shape gives the same results up to FP32
        summation order."""
        self._check(self._lib.oxdna_set_replica_tiling(self._ctx, int(workgroup), int(depth_divisor)),
                    "set_replica_tiling")
        self._cache.clear()

    def _upload_state(self):
        super()._upload_state()
        if self._offsets is None:
            raise RuntimeError("use create_state_from_snapshots for a replica ensemble")
        off = np.ascontiguousarray(self._offsets, dtype=np.uint32)
        box = np.ascontiguousarray(self._boxes, dtype=np.float64)
        self._check(self._lib.oxdna_set_replicas(self._ctx, len(off) - 1, off.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                 box.ctypes.data_as(C.POINTER(C.c_double))), "set_replicas")
        if self._bias is not None:
            self._upload_bias()

    @property
    def bias(self):
        return self._bias

    @bias.setter
    def bias(self, b):
        if b is not None and len(b.ends) != self.n_replicas:
            raise ValueError(f"bias has {len(b.ends)} replicas, the ensemble {self.n_replicas}")
        self._bias = b
        self._upload_bias()

    def _upload_bias(self):
        b = self._bias
        if b is None:
            self._check(self._lib.oxdna_set_replica_bias(self._ctx, None, None, None), "set_replica_bias")
        else:
            dp = C.POINTER(C.c_double)
            self._check(self._lib.oxdna_set_replica_bias(self._ctx, b.ends.ctypes.data_as(C.POINTER(C.c_int32)),
                                                         b.k.ctypes.data_as(dp), b.d0.ctypes.data_as(dp)),
                        "set_replica_bias")
        self._cache.clear()

    def replica_energies(self):
        """Energies of every replica at the current state: (n_replicas, 9) FP64, the
        split_energy slots then the bias (names in REPLICA_ENERGY_NAMES)."""
        if not self._attached:
            self._attach()
        key = "replica_energies"
        if key not in self._cache:
            self._check(self._lib.oxdna_compute(self._ctx, self._mask, 1), "compute")
            out = np.zeros((self.n_replicas, len(REPLICA_ENERGY_NAMES)))
            self._check(self._lib.oxdna_get_replica_energies(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double))),
                        "get_replica_energies")
            self._cache[key] = out
        return self._cache[key].copy()

    def replica_snapshot(self, r):
        """The current state of replica r as its own snapshot (its own box)."""
        s = self.state.get_snapshot()
        o0, o1 = int(self._offsets[r]), int(self._offsets[r + 1])
        out = io.Snapshot()
        out.particles.resize(o1 - o0)
        out.particles.types = list(s.particles.types)
        for name in ("position", "typeid", "mass", "moment_inertia", "orientation", "charge", "velocity", "angmom"):
            getattr(out.particles, name)[:] = getattr(s.particles, name)[o0:o1]
        g = np.asarray(s.bonds.group).reshape(-1, 2)
        keep = (g[:, 0] >= o0) & (g[:, 0] < o1)
        out.bonds.types = list(s.bonds.types)
        out.bonds.group = (g[keep] - o0).astype(np.int32)
        out.bonds.typeid = np.asarray(s.bonds.typeid)[keep].astype(np.int32)
        out.bonds.N = int(keep.sum())
        b = self._boxes[r]
        out.configuration.box = [b[0], b[1], b[2], 0.0, 0.0, 0.0]
        return out
