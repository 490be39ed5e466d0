"""Seeded synthetic systems: ideal duplexes, random duplex gases, hairpins.

Geometry of an ideal helix (oxDNA's generator convention): bases on successive
levels are rotated by ``twist`` about the helix axis and translated by ``rise``; a
nucleotide's a1 points from its centre of mass to the axis, its centre of mass sits
``com_radius`` (POS_BASE + 0.2 = 0.6) from the axis, and a3 runs along the axis in
the strand's 3'->5' direction. The complementary strand has a1' = -a1, a3' = -a3 at
the same level. Strands are listed 3'->5' (n3 = i-1, n5 = i+1), bonds are (i, n5[i]).
"""

import math

import numpy as np

from . import io

RISE = 0.3897
TWIST = 2.0 * math.pi / 10.34
COM_RADIUS = 0.6
COMPLEMENT = {"A": "T", "T": "A", "G": "C", "C": "G"}


class System:
    """Arrays of one system: pos (n,3), quat (n,4), typeid (n,), bonds (nb,2), box (3,),
    charge (n,), strand (n,), base letters, n3 / n5 neighbour indices."""

    def __init__(self, pos, quat, typeid, bonds, box, strand, bases, n3, n5, charge=None):
        self.pos = np.asarray(pos, dtype=float)
        self.quat = np.asarray(quat, dtype=float)
        self.typeid = np.asarray(typeid, dtype=np.int32)
        self.bonds = np.asarray(bonds, dtype=np.int32).reshape(-1, 2)
        self.box = np.asarray(box, dtype=float)
        self.strand = np.asarray(strand, dtype=int)
        self.bases = list(bases)
        self.n3 = np.asarray(n3, dtype=int)
        self.n5 = np.asarray(n5, dtype=int)
        self.charge = np.ones(len(self.pos)) if charge is None else np.asarray(charge, dtype=float)

    @property
    def n(self):
        return len(self.pos)

    def snapshot(self):
        return io.snapshot_from_arrays(self.pos, self.quat, self.typeid, self.box,
                                       bonds=[tuple(b) for b in self.bonds], charges=self.charge)

    def write(self, top_path, conf_path):
        io.write_topology(top_path, self.strand, self.bases, self.n3, self.n5)
        io.write_configuration(conf_path, 0, self.box, self.pos, self.quat)


def _axis_frame(k, twist):
    """a1 of strand-1 level k for a helix along +z: points to the axis."""
    phi = k * twist
    return np.array([-math.cos(phi), -math.sin(phi), 0.0])


def random_sequence(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(list(alphabet), size=n))


def ideal_duplex(seq, rise=RISE, twist=TWIST, com_radius=COM_RADIUS):
    """Both strands of an ideal B-form duplex along +z, centred on the origin.

    Returns (pos, a1, a3, bases, strand_id, n3, n5) with strand 0 = ``seq`` listed
    3'->5' and strand 1 its complement listed 3'->5'."""
    n = len(seq)
    z0 = -0.5 * (n - 1) * rise
    pos, a1s, a3s, bases = [], [], [], []
    for k in range(n):
        a1 = _axis_frame(k, twist)
        axis_pt = np.array([0.0, 0.0, z0 + k * rise])
        pos.append(axis_pt - com_radius * a1)
        a1s.append(a1)
        a3s.append(np.array([0.0, 0.0, 1.0]))
        bases.append(seq[k])
    for m in range(n):  # complement, 3'->5' runs down the axis
        k = n - 1 - m
        a1 = -_axis_frame(k, twist)
        axis_pt = np.array([0.0, 0.0, z0 + k * rise])
        pos.append(axis_pt - com_radius * a1)
        a1s.append(a1)
        a3s.append(np.array([0.0, 0.0, -1.0]))
        bases.append(COMPLEMENT[seq[k]])
    strand = [0] * n + [1] * n
    n3 = [i - 1 for i in range(n)] + [n + m - 1 if m > 0 else -1 for m in range(n)]
    n5 = [i + 1 if i < n - 1 else -1 for i in range(n)] + [n + m + 1 if m < n - 1 else -1 for m in range(n)]
    return np.array(pos), np.array(a1s), np.array(a3s), bases, strand, n3, n5


def duplex_system(seq, box=20.0, half_charged_ends=False):
    """One ideal duplex in a cubic box (box-centred coordinates)."""
    pos, a1, a3, bases, strand, n3, n5 = ideal_duplex(seq)
    quat = io.orientations_to_quaternions(a1, a3)
    typeid = [io.base_typeid(b) for b in bases]
    bonds = [(i, n5[i]) for i in range(len(bases)) if n5[i] >= 0]
    top = dict(n=len(bases), n3=np.array(n3), n5=np.array(n5))
    charge = io.terminal_charges(top) if half_charged_ends else None
    return System(pos, quat, typeid, bonds, [box] * 3, strand, bases, n3, n5, charge)


def haar_rotations(rng, n):
    """n uniformly random rotation matrices (normalised Gaussian quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _min_image(d, L):
    return d - L * np.rint(d / L)


def duplex_gas(n_duplexes, n_bp, box, seed=0, min_dist=2.0, max_tries=200000):
    """``n_duplexes`` ideal duplexes of ``n_bp`` bp with seeded random sequences,
    Haar-random orientations and uniform centres in a cubic box of side ``box``,
    rejecting placements with any inter-duplex nucleotide pair closer than min_dist."""
    rng = np.random.default_rng(seed)
    L = float(box)
    half_len = 0.5 * n_bp * RISE + COM_RADIUS + 0.5
    reach = 2 * half_len + min_dist
    ncell = max(1, int(L // reach))
    cell = L / ncell
    grid = {}
    placed_pos = []
    all_pos, all_q, all_t, all_b, strands, bases, n3s, n5s = [], [], [], [], [], [], [], []
    offset = 0
    tries = 0
    while len(placed_pos) < n_duplexes:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place all duplexes without overlap")
        seq = random_sequence(rng, n_bp)
        pos, a1, a3, bs, strand, n3, n5 = ideal_duplex(seq)
        R = haar_rotations(rng, 1)[0]
        centre = rng.uniform(-L / 2, L / 2, size=3)
        p = pos @ R.T + centre
        ci = tuple((np.floor((centre + L / 2) / cell).astype(int)) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((ci[0] + dx) % ncell, (ci[1] + dy) % ncell, (ci[2] + dz) % ncell)
                    for k in grid.get(key, ()):
                        d = _min_image(p[:, None, :] - placed_pos[k][None, :, :], L)
                        if np.min(np.einsum("ijk,ijk->ij", d, d)) < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        grid.setdefault(ci, []).append(len(placed_pos))
        placed_pos.append(p)
        p = _min_image(p, L)
        all_pos.append(p)
        all_q.append(io.orientations_to_quaternions(a1 @ R.T, a3 @ R.T, check=False))
        all_t.extend(io.base_typeid(b) for b in bs)
        nd = len(bs)
        all_b.extend((i + offset, n5[i] + offset) for i in range(nd) if n5[i] >= 0)
        strands.extend(s + 2 * (len(placed_pos) - 1) for s in strand)
        bases.extend(bs)
        n3s.extend(x + offset if x >= 0 else -1 for x in n3)
        n5s.extend(x + offset if x >= 0 else -1 for x in n5)
        offset += nd
    return System(np.concatenate(all_pos), np.concatenate(all_q), all_t, all_b, [L] * 3,
                  strands, bases, n3s, n5s)


def hairpin(stem_seq, loop_len, loop_base="T", rise=RISE, twist=TWIST, com_radius=COM_RADIUS):
    """One hairpin: stem ``stem_seq`` (5' half, listed 3'->5'), a poly-``loop_base`` loop of
    ``loop_len`` nt, then the complementary 3' half; one strand, centred on the origin.

    The loop nucleotides sit on a circular arc above the stem that joins the two stem
    ends, with a1 pointing to the arc's centre and a3 along the arc."""
    s = len(stem_seq)
    pos, a1, a3, _, _, _, _ = ideal_duplex(stem_seq, rise, twist, com_radius)
    up_p, up_a1, up_a3 = pos[:s], a1[:s], a3[:s]
    dn_p, dn_a1, dn_a3 = pos[s:], a1[s:], a3[s:]
    start, end = up_p[-1], dn_p[0]
    mid = 0.5 * (start + end)
    chord = end - start
    c = np.linalg.norm(chord)
    u = chord / c
    w = np.array([0.0, 0.0, 1.0])
    w = w - u * np.dot(w, u)
    w /= np.linalg.norm(w)
    # arc of loop_len + 1 equal steps of length ~0.7 from start to end, bulging along +z
    step = 0.7
    n_seg = loop_len + 1
    arc_len = step * n_seg
    # solve for circle radius r with arc length arc_len over the chord c (theta in (0, 2pi))
    lo, hi = 1e-6, 2 * math.pi - 1e-6
    for _ in range(200):
        th = 0.5 * (lo + hi)
        if (c / (2 * math.sin(th / 2))) * th < arc_len:
            lo = th
        else:
            hi = th
    th = 0.5 * (lo + hi)
    r = c / (2 * math.sin(th / 2))
    h = r * math.cos(th / 2)  # centre offset from the chord midpoint (towards -w when th > pi)
    centre = mid - w * h
    loop_p, loop_a1, loop_a3 = [], [], []
    v0 = start - centre
    for k in range(1, loop_len + 1):
        ang = th * k / n_seg
        # rotate v0 about the axis n = u x w ... in the plane (u, w) from start to end
        e1 = v0 / np.linalg.norm(v0)
        e2 = u - e1 * np.dot(u, e1)
        if np.linalg.norm(e2) < 1e-9:
            e2 = w
        e2 /= np.linalg.norm(e2)
        p = centre + r * (math.cos(ang) * e1 + math.sin(ang) * e2)
        radial = (centre - p) / r
        tang = -math.sin(ang) * e1 + math.cos(ang) * e2
        a1k = radial
        a3k = tang - a1k * np.dot(tang, a1k)
        a3k /= np.linalg.norm(a3k)
        loop_p.append(p); loop_a1.append(a1k); loop_a3.append(a3k)
    P = np.concatenate([up_p, np.array(loop_p).reshape(-1, 3), dn_p])
    A1 = np.concatenate([up_a1, np.array(loop_a1).reshape(-1, 3), dn_a1])
    A3 = np.concatenate([up_a3, np.array(loop_a3).reshape(-1, 3), dn_a3])
    bases = list(stem_seq) + [loop_base] * loop_len + [COMPLEMENT[b] for b in reversed(stem_seq)]
    n = len(bases)
    n3 = [i - 1 for i in range(n)]
    n5 = [i + 1 if i < n - 1 else -1 for i in range(n)]
    return P - P.mean(axis=0), A1, A3, bases, [0] * n, n3, n5


def hairpin_ensemble(n_replicas, seed=3, stem=(4, 30), loop=(4, 10), box=30.0):
    """Replicas of seeded random hairpins (stem s ~ U{stem}, loop l ~ U{loop} poly-T),
    each centred in its own cubic box of side ``box``. Returns a list of System."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_replicas):
        s = int(rng.integers(stem[0], stem[1] + 1))
        ell = int(rng.integers(loop[0], loop[1] + 1))
        seq = random_sequence(rng, s)
        p, a1, a3, bases, strand, n3, n5 = hairpin(seq, ell)
        q = io.orientations_to_quaternions(a1, a3, check=False)
        bonds = [(i, n5[i]) for i in range(len(bases)) if n5[i] >= 0]
        out.append(System(p, q, [io.base_typeid(b) for b in bases], bonds, [box] * 3, strand,
                          bases, n3, n5))
    return out
