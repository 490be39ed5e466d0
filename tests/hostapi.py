"""Writes system files for tests/cpp/host_api_check and reads its output."""

import ctypes as C
import os
import subprocess

import numpy as np

from oxdna import io

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "host_api_check")


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    return EXE


def write_system(path, system, forces, types=tuple(io.DNA_BASES)):
    """The system and the packed parameters of `forces` in host_api_check's format."""
    n = system.n
    bonds = np.asarray(system.bonds, dtype=np.int32).reshape(-1, 2)
    with open(path, "wb") as fh:
        fh.write(np.array([n, len(bonds)], dtype=np.uint32).tobytes())
        fh.write(np.ascontiguousarray(system.pos, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(system.quat, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(system.typeid, dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(system.charge, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(system.box, dtype=np.float64).tobytes())
        fh.write(np.column_stack([bonds, np.zeros(len(bonds), dtype=np.int32)]).astype(np.int32).tobytes())
        by_term = {f._term: f for f in forces}
        for t in range(6):
            f = by_term.get(t)
            fh.write(np.array([1 if f is not None else 0], dtype=np.int32).tobytes())
            if f is not None:
                table, rcut = f._pack(list(types))
                fh.write(bytes(table))
                fh.write(np.asarray(list(rcut), dtype=np.float64).tobytes())
        b = by_term.get(6)
        fh.write(np.array([1 if b is not None else 0], dtype=np.int32).tobytes())
        if b is not None:
            fh.write(bytes(b._pack(["backbone"])[0]))


def read_output(path, n):
    """{term: (energies[8], force[n,4], torque[n,3])} in the order written."""
    out = {}
    raw = open(path, "rb").read()
    off = 0
    while off < len(raw):
        term = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
        off += 4
        e = np.frombuffer(raw, dtype=np.float64, count=8, offset=off).copy()
        off += 64
        f = np.frombuffer(raw, dtype=np.float64, count=4 * n, offset=off).reshape(n, 4).copy()
        off += 32 * n
        t = np.frombuffer(raw, dtype=np.float64, count=3 * n, offset=off).reshape(n, 3).copy()
        off += 24 * n
        out[term] = (e, f, t)
    return out
