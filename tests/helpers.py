"""Shared test helpers: synthetic systems, finite differences, tolerances."""

import numpy as np

from oxdna import io, systems


def quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def delta_quat(axis, eps):
    return np.array([np.cos(eps / 2.0), *(np.sin(eps / 2.0) * np.asarray(axis))])


def axes_quat(a1, a3):
    return io.orientations_to_quaternions(np.atleast_2d(a1), np.atleast_2d(a3))[0]


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def fd_forces_torques(energy, pos, quat, idx, h=1e-6, eps=1e-6):
    """Central differences of energy(pos, quat) at particle idx: (-dE/dx, -dE/dtheta)
    with left-multiplied rotations (lab-frame torque), the reference's fd_check method
    (reference pytest/_helpers.py:137-181)."""
    f = np.zeros(3)
    t = np.zeros(3)
    for d in range(3):
        pp = pos.copy(); pp[idx, d] += h
        ep = energy(pp, quat)
        pp[idx, d] -= 2 * h
        em = energy(pp, quat)
        f[d] = -(ep - em) / (2 * h)
    for k, axis in enumerate(np.eye(3)):
        qp = quat.copy(); qp[idx] = quat_mul(delta_quat(axis, eps), quat[idx])
        ep = energy(pos, qp)
        qm = quat.copy(); qm[idx] = quat_mul(delta_quat(axis, -eps), quat[idx])
        em = energy(pos, qm)
        t[k] = -(ep - em) / (2 * eps)
    return f, t


def cfg1_system(seed=0):
    """One ideal 20-bp B-duplex (40 nt) with a seeded random sequence in a box of 20."""
    rng = np.random.default_rng(seed)
    return systems.duplex_system(systems.random_sequence(rng, 20), box=20.0)


def jiggle(system, seed=1, dx=0.05, dq=0.05):
    """Copy of a system with small random displacements / rotations (thermal-like)."""
    rng = np.random.default_rng(seed)
    s = systems.System(system.pos + rng.normal(scale=dx, size=system.pos.shape), system.quat.copy(),
                       system.typeid, system.bonds, system.box, system.strand, system.bases, system.n3,
                       system.n5, system.charge)
    for i in range(s.n):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        s.quat[i] = quat_mul(delta_quat(ax, rng.normal(scale=dq)), s.quat[i])
    return s
