"""oxdna_pair_forces_noalloc (library-held fixed scratch, no device allocation) against
oxdna_pair_forces (scratch the library allocates): the same kernels on the same inputs,
so every output must agree bit for bit. Two calls in a row on one device, the second on
a smaller system, so the second reads scratch the first left behind; and a system past
the fixed maximum is rejected with an error before anything is launched."""

import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H
from oxdna import systems
from oxdna.model import dna2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(system):
    from hipmem import DeviceArray
    n = system.n
    pos4 = np.zeros((n, 4))
    pos4[:, :3] = system.pos
    pos4[:, 3] = np.frombuffer(np.asarray(system.typeid, dtype=np.int64).tobytes(), dtype=np.float64)
    return DeviceArray(pos4), DeviceArray(system.quat), DeviceArray(system.charge)


def _call(lib, entry, f, system, dev, nlist, table, rcutsq):
    from hipmem import DeviceArray
    from oxdna import _abi
    n = system.n
    d_pos, d_q, d_c = dev
    nn, rows, head = nlist
    box = _abi.Box()
    box.L[0], box.L[1], box.L[2] = system.box
    out_f = DeviceArray(shape=(n, 4), dtype=np.float32)
    out_t = DeviceArray(shape=(n, 4), dtype=np.float32)
    out_v = DeviceArray(shape=(6, n), dtype=np.float32)
    out_e = DeviceArray(shape=(8,), dtype=np.float64)
    rc = getattr(lib, entry)(f._term, out_f.ptr, out_t.ptr, out_v.ptr, n, n, d_pos.ptr, d_q.ptr, d_c.ptr,
                             C.byref(box), nn.ptr, rows.ptr, head.ptr, 4, C.cast(table, C.c_void_p), rcutsq,
                             out_e.ptr, None)
    assert rc == 0, lib.oxdna_last_error(None)
    return out_f.get(), out_t.get(), out_v.get(), out_e.get()


def _compare_on(lib, system, label):
    from test_gpu_parity import _device_nlist
    forces, _ = dna2.forces(fused=True)
    dev = _inputs(system)
    for f in forces[1:]:
        table, rcut = f._pack(["A", "G", "C", "T"])
        rcutsq = (C.c_double * 16)(*[r * r for r in rcut])
        nlist = _device_nlist(system, max(rcut) + 0.4)
        want = _call(lib, "oxdna_pair_forces", f, system, dev, nlist, table, rcutsq)
        got = _call(lib, "oxdna_pair_forces_noalloc", f, system, dev, nlist, table, rcutsq)
        _say(f"[{label}] {type(f).__name__}: n={system.n} E={np.round(want[3], 6).tolist()} "
             f"max|F|={np.abs(want[0][:, :3]).max():.3e}")
        for name, a, b in zip(("force", "torque", "virial", "energy"), want, got):
            assert np.array_equal(a, b), f"{label} {type(f).__name__}: {name} differs (max {np.abs(a - b).max()})"


def test_noalloc_matches_allocating_driver_twice():
    from oxdna import _engine
    lib = _engine.load()
    big = H.jiggle(systems.duplex_gas(8, 20, box=40.0, seed=4), seed=4, dx=0.03, dq=0.05)
    small = H.jiggle(H.cfg1_system(7), seed=7, dx=0.03, dq=0.05)
    _compare_on(lib, big, "first call, 320 nt")
    _compare_on(lib, small, "second call, 40 nt")


def test_noalloc_rejects_past_its_maximum():
    from oxdna import _abi, _engine
    lib = _engine.load()
    forces, _ = dna2.forces(fused=True)
    table, rcut = forces[1]._pack(["A", "G", "C", "T"])
    rcutsq = (C.c_double * 16)(*[r * r for r in rcut])
    box = _abi.Box()
    box.L[0] = box.L[1] = box.L[2] = 100.0
    fake = C.c_void_p(256)  # never dereferenced: the size check precedes every launch
    rc = lib.oxdna_pair_forces_noalloc(forces[1]._term, fake, fake, None, 0, 65537, fake, fake, fake, C.byref(box),
                                       fake, fake, fake, 4, C.cast(table, C.c_void_p), rcutsq, None, None)
    msg = lib.oxdna_last_error(None).decode()
    _say("rejection:", rc, msg)
    assert rc != 0 and "65536" in msg
