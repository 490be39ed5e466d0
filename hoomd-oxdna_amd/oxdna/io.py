"""oxDNA topology / configuration I/O and system snapshots.

Conventions as the reference (reference oxdna/io.py): base letters map to type ids
A=0, G=1, C=2, T=3 (U shares T's id); a nucleotide's orientation is the quaternion
(w, x, y, z) of the rotation whose columns are (a1, a2 = a3 x a1, a3), so that
rotate(q, e_x) = a1 and rotate(q, e_z) = a3; positions are centred on the box
(oxDNA writes them in [0, L]); bonds are (i, n5[i]); charges default to 1.0 and
oxDNA2 / oxRNA2 put 0.5 at strand ends when asked.

Adds writers for both formats so systems built here load into oxDNA itself.
"""

import numpy as np

DNA_BASES = ["A", "G", "C", "T"]
BASE_ALIAS = {"U": "T"}


def base_typeid(letter):
    """Type id of a base letter (U counts as T)."""
    return DNA_BASES.index(BASE_ALIAS.get(letter, letter))


# --------------------------------------------------------------------------- files
def read_topology(top_path):
    """Parse an oxDNA .top file into ``dict(n, n_strands, strand, base, n3, n5)``
    (n3 / n5: 3' / 5' neighbour index, -1 if none)."""
    with open(top_path) as fh:
        head = fh.readline().split()
        n, ns = int(head[0]), int(head[1])
        rows = [fh.readline().split() for _ in range(n)]
    return dict(n=n, n_strands=ns,
                strand=np.array([int(r[0]) for r in rows], dtype=int),
                base=[r[1] for r in rows],
                n3=np.array([int(r[2]) for r in rows], dtype=int),
                n5=np.array([int(r[3]) for r in rows], dtype=int))


def iter_conf_frames(conf_path):
    """Yield ``(t, box, rows)`` per frame of an oxDNA configuration / trajectory;
    rows are (n, 15): com, a1, a3, velocity, angular velocity."""
    frame_t, frame_box, rows = None, None, []
    with open(conf_path) as fh:
        lines = iter(fh.readlines())
    for line in lines:
        if line.startswith("t ="):
            if frame_t is not None:
                yield frame_t, frame_box, np.array(rows, dtype=float)
            frame_t = float(line.split("=")[1])
            frame_box = np.array(next(lines).split("=")[1].split(), dtype=float)
            next(lines)  # energy line
            rows = []
            continue
        parts = line.split()
        if frame_t is not None and len(parts) >= 9:
            vals = [float(x) for x in parts[:15]]
            rows.append(vals + [0.0] * (15 - len(vals)))
    if frame_t is not None:
        yield frame_t, frame_box, np.array(rows, dtype=float)


def write_topology(path, strand, bases, n3, n5):
    with open(path, "w") as fh:
        fh.write(f"{len(bases)} {len(set(int(s) for s in strand))}\n")
        for s, b, a, c in zip(strand, bases, n3, n5):
            fh.write(f"{int(s)} {b} {int(a)} {int(c)}\n")


def write_configuration(path, t, box, pos, quat, vel=None, angvel=None, energy=(0.0, 0.0, 0.0),
                        append=False):
    """Append or write one oxDNA frame; positions are shifted back into [0, L]."""
    box = np.asarray(box, dtype=float)
    a1 = rotate(quat, np.array([1.0, 0.0, 0.0]))
    a3 = rotate(quat, np.array([0.0, 0.0, 1.0]))
    n = len(pos)
    vel = np.zeros((n, 3)) if vel is None else vel
    angvel = np.zeros((n, 3)) if angvel is None else angvel
    with open(path, "a" if append else "w") as fh:
        fh.write(f"t = {t}\nb = {box[0]} {box[1]} {box[2]}\nE = {energy[0]} {energy[1]} {energy[2]}\n")
        p = np.asarray(pos) + box / 2.0
        for k in range(n):
            cols = np.concatenate([p[k], a1[k], a3[k], vel[k], angvel[k]])
            fh.write(" ".join(f"{x:.15g}" for x in cols) + "\n")


def body_angular_momentum(quat, angmom):
    """L in the body frame from the quaternion momentum p of the engine's state
    (HOOMD's convention: q* (x) p = (0, 2 L_body))."""
    q = np.asarray(quat, dtype=float)
    p = np.asarray(angmom, dtype=float)
    qw, qv = q[..., :1], -q[..., 1:]  # q*
    pw, pv = p[..., :1], p[..., 1:]
    return 0.5 * (qw * pv + pw * qv + np.cross(qv, pv))


def angular_velocity(quat, angmom, moment_inertia):
    """Lab-frame angular velocity R(q) (L_body / I) per particle."""
    return rotate(quat, body_angular_momentum(quat, angmom) / np.asarray(moment_inertia, dtype=float))


def kinetic_energy(mass, velocity, quat, angmom, moment_inertia):
    """Translational plus rotational kinetic energy of the whole system."""
    lb = body_angular_momentum(quat, angmom)
    m = np.asarray(mass, dtype=float)
    v = np.asarray(velocity, dtype=float)
    return float(0.5 * np.sum(m[:, None] * v * v) + 0.5 * np.sum(lb * lb / np.asarray(moment_inertia, dtype=float)))


# ------------------------------------------------------------------- quaternions
def rotate(q, v):
    """Rotate v (3,) or (n,3) by unit quaternions q (n,4) (w,x,y,z); returns (n,3)."""
    q = np.atleast_2d(np.asarray(q, dtype=float))
    v = np.broadcast_to(np.asarray(v, dtype=float), (q.shape[0], 3))
    s, u = q[:, :1], q[:, 1:]
    return ((s * s - np.sum(u * u, axis=1, keepdims=True)) * v
            + 2.0 * s * np.cross(u, v) + 2.0 * np.sum(u * v, axis=1, keepdims=True) * u)


def frames_to_quaternions(R):
    """Rotation matrices (n,3,3) -> unit quaternions (n,4) (w,x,y,z), w >= 0.

    Branch on the largest of (trace, R00, R11, R22) so the square root is taken of
    the largest of 4w^2, 4x^2, 4y^2, 4z^2 (Shepperd's stable selection)."""
    R = np.asarray(R, dtype=float)
    n = R.shape[0]
    q = np.empty((n, 4))
    tr = R[:, 0, 0] + R[:, 1, 1] + R[:, 2, 2]
    cand = np.stack([tr, R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]], axis=1)
    pick = np.argmax(cand, axis=1)
    for k in range(n):
        m = R[k]
        c = pick[k]
        if c == 0:
            w = 0.5 * np.sqrt(1.0 + tr[k])
            q[k] = (w, (m[2, 1] - m[1, 2]) / (4 * w), (m[0, 2] - m[2, 0]) / (4 * w),
                    (m[1, 0] - m[0, 1]) / (4 * w))
        else:
            i = c - 1
            j, l = (i + 1) % 3, (i + 2) % 3
            big = 0.5 * np.sqrt(1.0 + m[i, i] - m[j, j] - m[l, l])
            out = np.empty(4)
            out[1 + i] = big
            out[0] = (m[l, j] - m[j, l]) / (4 * big)
            out[1 + j] = (m[j, i] + m[i, j]) / (4 * big)
            out[1 + l] = (m[l, i] + m[i, l]) / (4 * big)
            q[k] = out
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1.0
    return q


def orientations_to_quaternions(a1, a3, check=True):
    """Per-nucleotide (a1, a3) -> quaternions with rotate(q, e_x) = a1, rotate(q, e_z) = a3."""
    e1 = np.asarray(a1, dtype=float)
    e3 = np.asarray(a3, dtype=float)
    e1 = e1 / np.linalg.norm(e1, axis=1, keepdims=True)
    e3 = e3 / np.linalg.norm(e3, axis=1, keepdims=True)
    e2 = np.cross(e3, e1)
    R = np.stack([e1, e2, e3], axis=2)  # columns a1, a2, a3
    q = frames_to_quaternions(R)
    if check:
        err = np.linalg.norm(rotate(q, [1.0, 0.0, 0.0]) - e1, axis=1)
        if err.size and err.max() > 1e-6:
            i = int(np.argmax(err))
            raise ValueError(f"quaternion convention check failed at nucleotide {i}: "
                             f"|rotate(q,ex) - a1| = {err[i]:.2e}")
    return q


# -------------------------------------------------------------------- snapshots
class _Particles:
    def __init__(self):
        self.N = 0
        self.types = list(DNA_BASES)
        self.position = np.zeros((0, 3))
        self.typeid = np.zeros(0, dtype=np.int32)
        self.mass = np.zeros(0)
        self.moment_inertia = np.zeros((0, 3))
        self.orientation = np.zeros((0, 4))
        self.charge = np.zeros(0)
        self.velocity = np.zeros((0, 3))
        self.angmom = np.zeros((0, 4))

    def resize(self, n):
        self.N = n
        self.position = np.zeros((n, 3))
        self.typeid = np.zeros(n, dtype=np.int32)
        self.mass = np.ones(n)
        self.moment_inertia = np.ones((n, 3))
        self.orientation = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
        self.charge = np.ones(n)
        self.velocity = np.zeros((n, 3))
        self.angmom = np.zeros((n, 4))


class _Bonds:
    def __init__(self):
        self.N = 0
        self.types = ["backbone"]
        self.group = np.zeros((0, 2), dtype=np.int32)
        self.typeid = np.zeros(0, dtype=np.int32)


class _Configuration:
    def __init__(self):
        self.box = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]


class Snapshot:
    """System snapshot with the attribute layout of the host engine's snapshot that
    the reference builds (particles / bonds / configuration)."""

    def __init__(self):
        self.particles = _Particles()
        self.bonds = _Bonds()
        self.configuration = _Configuration()

    def copy(self):
        s = Snapshot()
        for part in ("particles", "bonds", "configuration"):
            src, dst = getattr(self, part), getattr(s, part)
            for k, v in vars(src).items():
                setattr(dst, k, v.copy() if hasattr(v, "copy") else v)
        return s


def snapshot_from_arrays(positions, quaternions, typeids, box, bonds=None, types=DNA_BASES,
                         bond_type="backbone", charges=None):
    """Snapshot from arrays (positions already box-centred); unit charges unless given."""
    n = len(positions)
    snap = Snapshot()
    snap.particles.resize(n)
    snap.particles.types = list(types)
    snap.particles.position[:] = positions
    snap.particles.typeid[:] = typeids
    snap.particles.orientation[:] = quaternions
    snap.particles.charge[:] = np.ones(n) if charges is None else charges
    b = np.asarray(box, dtype=float).ravel()
    snap.configuration.box = [b[0], b[1], b[2], 0.0, 0.0, 0.0]
    bonds = [] if bonds is None else list(bonds)
    snap.bonds.N = len(bonds)
    snap.bonds.types = [bond_type]
    snap.bonds.group = np.array(bonds, dtype=np.int32).reshape(-1, 2)
    snap.bonds.typeid = np.zeros(len(bonds), dtype=np.int32)
    return snap


def terminal_charges(top, end_charge=0.5, interior_charge=1.0):
    """Per-nucleotide charge with ``end_charge`` where the 3' or 5' neighbour is missing."""
    q = np.full(top["n"], interior_charge, dtype=float)
    q[(np.asarray(top["n3"]) < 0) | (np.asarray(top["n5"]) < 0)] = end_charge
    return q


def build_snapshot(top_path, conf_path, frame=0, bond_type="backbone", half_charged_ends=False):
    """Snapshot from an oxDNA topology + one configuration frame."""
    top = read_topology(top_path)
    frames = list(iter_conf_frames(conf_path))
    _, box, rows = frames[frame]
    pos = rows[:, 0:3] - box / 2.0
    quats = orientations_to_quaternions(rows[:, 3:6], rows[:, 6:9])
    typeids = np.array([base_typeid(b) for b in top["base"]], dtype=np.int32)
    bonds = [(i, int(top["n5"][i])) for i in range(top["n"]) if top["n5"][i] >= 0]
    charges = terminal_charges(top) if half_charged_ends else np.ones(top["n"])
    return snapshot_from_arrays(pos, quats, typeids, box, bonds=bonds, bond_type=bond_type,
                                charges=charges)
