"""GPU: the stateless drivers hold no shared device state.

Two host threads, each with its own HIP stream, call oxdna_pair_forces (fused
HB + cross-stacking, with virial and energies) and oxdna_bonded_forces on two
different systems at the same time, many times over. Every concurrent result must
equal, bit for bit, what one call of the same driver on the same inputs gives
alone (the kernels and their summation order do not depend on what else runs).
Before each call's workspace was its own allocation, the two threads' frames and
parameter blocks overwrote each other.
"""

import ctypes as C
import sys
import threading

import numpy as np
import pytest

import helpers as H
import oracle as O
from oxdna import _abi, _engine, systems
from oxdna.model import dna2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    from hipmem import hip
    L = hip()
    L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipStreamSynchronize.argtypes = [C.c_void_p]
    L.hipStreamDestroy.argtypes = [C.c_void_p]
    L.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return L


class _Case:
    """Device inputs and outputs of one system for both drivers."""

    def __init__(self, system):
        from hipmem import DeviceArray
        self.n = n = system.n
        pos4 = np.zeros((n, 4))
        pos4[:, :3] = system.pos
        pos4[:, 3] = np.frombuffer(np.asarray(system.typeid, dtype=np.int64).tobytes(), dtype=np.float64)
        self.pos, self.q, self.c = DeviceArray(pos4), DeviceArray(system.quat), DeviceArray(system.charge)
        self.box = _abi.Box()
        self.box.L[0], self.box.L[1], self.box.L[2] = system.box
        forces, _ = dna2.forces(fused=True)
        f = [x for x in forces if x._term == _abi.TERM_HBONDCROSS][0]
        self.term = f._term
        self.table, rcut = f._pack(["A", "G", "C", "T"])
        self.rcutsq = (C.c_double * 16)(*[r * r for r in rcut])
        pairs = O.Oracle(system.pos, system.quat, system.typeid, system.box, bonds=system.bonds).pairs(
            max(rcut) + 0.4, False)
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
        self.nn, self.nl = DeviceArray(nn), DeviceArray(nl)
        self.head = DeviceArray(np.arange(n, dtype=np.uint64) * width)
        self.bparams = forces[0]._pack(["backbone"])
        blist = np.zeros((2, n, 2), dtype=np.uint32)
        bpos = np.zeros((2, n), dtype=np.uint32)
        nb = np.zeros(n, dtype=np.uint32)
        for a, b in system.bonds:
            for me, other, role in ((a, b, 0), (b, a, 1)):
                blist[nb[me], me] = (other, 0)
                bpos[nb[me], me] = role
                nb[me] += 1
        self.bl, self.bp, self.nb = DeviceArray(blist), DeviceArray(bpos), DeviceArray(nb)
        self.out = [DeviceArray(shape=(n, 4), dtype=np.float32), DeviceArray(shape=(n, 4), dtype=np.float32),
                    DeviceArray(shape=(6, n), dtype=np.float32), DeviceArray(shape=(8,), dtype=np.float64),
                    DeviceArray(shape=(n, 4), dtype=np.float32), DeviceArray(shape=(n, 4), dtype=np.float32),
                    DeviceArray(shape=(n, 4), dtype=np.float64)]

    def call(self, lib, stream):
        f, t, v, e, bf, bt, split = self.out
        rc = lib.oxdna_pair_forces(self.term, f.ptr, t.ptr, v.ptr, self.n, self.n, self.pos.ptr, self.q.ptr,
                                   self.c.ptr, C.byref(self.box), self.nn.ptr, self.nl.ptr, self.head.ptr, 4,
                                   C.cast(self.table, C.c_void_p), self.rcutsq, e.ptr, stream)
        assert rc == 0, lib.oxdna_last_error(None)
        rc = lib.oxdna_bonded_forces(bf.ptr, bt.ptr, None, 0, split.ptr, self.n, self.pos.ptr, self.q.ptr,
                                     C.byref(self.box), self.bl.ptr, self.bp.ptr, self.n, self.nb.ptr,
                                     self.bparams, 1, stream)
        assert rc == 0, lib.oxdna_last_error(None)

    def read(self, hip, stream):
        outs = []
        for d in self.out:
            h = np.empty(d.shape, dtype=d.dtype)
            assert hip.hipMemcpyAsync(h.ctypes.data, d.ptr, d.nbytes, 2, stream) == 0
            outs.append(h)
        assert hip.hipStreamSynchronize(stream) == 0
        return outs


def test_concurrent_driver_calls_are_bit_identical_to_lone_calls():
    lib = _engine.load()
    hip = _hip()
    cases = [_Case(H.jiggle(H.cfg1_system(3), seed=3)), _Case(H.jiggle(systems.duplex_gas(12, 20, box=40.0, seed=6),
                                                                        seed=6))]
    _say("systems:", [c.n for c in cases])
    ref = []
    for c in cases:
        c.call(lib, None)
        ref.append(c.read(hip, None))
    errors = []

    def worker(c, r, iters=40):
        s = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(s)) == 0
        try:
            for it in range(iters):
                c.call(lib, s)
                got = c.read(hip, s)
                for k, (g, w) in enumerate(zip(got, r)):
                    if not np.array_equal(g, w):
                        errors.append((c.n, it, k, float(np.abs(g.astype(float) - w.astype(float)).max())))
                        return
        except Exception as exc:  # reported by the main thread
            errors.append((c.n, "exception", repr(exc)))
        finally:
            hip.hipStreamSynchronize(s)
            hip.hipStreamDestroy(s)

    threads = [threading.Thread(target=worker, args=(c, r)) for c, r in zip(cases, ref)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a driver thread did not finish"
    assert not errors, errors
    _say("40 concurrent call pairs per thread, every output bit-identical to the lone call")
