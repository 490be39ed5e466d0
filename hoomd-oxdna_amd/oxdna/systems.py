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


# backbone site = centre + c1 a1 + c2 a2 + c3 a3 with a2 = a3 x a1
DNA1_BACK = (-0.4, 0.0, 0.0)        # oxDNA1
DNA2_BACK = (-0.34, 0.3408, 0.0)    # oxDNA2 (grooved)
RNA_BACK = (-0.4, 0.0, 0.2)         # oxRNA


def _bridge(first, last, m, bulge, back=DNA2_BACK, step=0.75):
    """m single-stranded nucleotides joining nucleotide ``first`` to ``last`` (each a
    (centre, a1, a3) triple). Their backbone sites (centre + back . (a1, a3 x a1, a3))
    lie on the straight segment between the anchors' backbone sites when that is at
    least (m + 1) * step long, otherwise on a circular arc bulging along ``bulge``
    whose consecutive backbone sites, anchors included, are exactly ``step`` apart.
    a3 follows the chain, a1 is normal to the arc's plane (or across the segment), so
    the base sites run parallel to the backbone sites and never crowd.
    Returns centres, a1, a3 (m rows each)."""
    c1, c2, c3 = back

    def site(p, a1, a3):
        p, a1, a3 = (np.asarray(v, float) for v in (p, a1, a3))
        return p + c1 * a1 + c2 * np.cross(a3, a1) + c3 * a3

    bs = site(*first)
    be = site(*last)
    if m == 0:
        return np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3))
    chord = be - bs
    c = np.linalg.norm(chord)
    u = chord / c
    w = np.asarray(bulge, float) - u * np.dot(bulge, u)
    if np.linalg.norm(w) < 1e-9:
        w = np.cross(u, [0.0, 0.0, 1.0]) if abs(u[2]) < 0.9 else np.cross(u, [1.0, 0.0, 0.0])
    w /= np.linalg.norm(w)
    B, A1, A3 = [], [], []
    if c >= step * (m + 1):
        for k in range(1, m + 1):
            B.append(bs + chord * k / (m + 1))
            A1.append(-w)
            A3.append(u)
    else:
        # equal chords `step` on a circle: sin((m+1) x / 2) / sin(x / 2) = c / step for
        # the angle x per step (the ratio falls from m + 1 to 0 on (0, 2 pi / (m + 1)))
        lo, hi = 1e-9, 2 * math.pi / (m + 1) - 1e-9
        for _ in range(200):
            x = 0.5 * (lo + hi)
            if math.sin((m + 1) * x / 2) / math.sin(x / 2) > c / step:
                lo = x
            else:
                hi = x
        x = 0.5 * (lo + hi)
        r = step / (2 * math.sin(x / 2))
        th = (m + 1) * x
        h = r * math.cos(th / 2)
        centre = 0.5 * (bs + be) - w * h
        e1 = (bs - centre) / r
        e2 = u - e1 * np.dot(u, e1)
        e2 /= np.linalg.norm(e2)
        # rotate from the first anchor through the bulge to the last: pick the sense
        # that lands on it
        if np.linalg.norm(centre + r * (math.cos(th) * e1 + math.sin(th) * e2) - be) > 1e-6 * max(1.0, r):
            e2 = -e2
        normal = np.cross(e1, e2)
        for k in range(1, m + 1):
            ang = x * k
            b = centre + r * (math.cos(ang) * e1 + math.sin(ang) * e2)
            tang = -math.sin(ang) * e1 + math.cos(ang) * e2
            B.append(b); A1.append(normal); A3.append(tang)
    B, A1, A3 = np.array(B), np.array(A1), np.array(A3)
    return B - c1 * A1 - c2 * np.cross(A3, A1) - c3 * A3, A1, A3


def hairpin(stem_seq, loop_len, loop_base="T", rise=RISE, twist=TWIST, com_radius=COM_RADIUS, back=DNA2_BACK):
    """One hairpin: stem ``stem_seq`` (5' half, listed 3'->5'), a poly-``loop_base`` loop of
    ``loop_len`` nt, then the complementary 3' half; one strand, centred on the origin.

    The loop's backbone sites sit on a circular arc above the stem, spaced like the
    FENE bond length (backbone site ``back``: oxDNA2 by default), with a3 along the
    arc and a1 normal to its plane."""
    s = len(stem_seq)
    pos, a1, a3, _, _, _, _ = ideal_duplex(stem_seq, rise, twist, com_radius)
    up_p, up_a1, up_a3 = pos[:s], a1[:s], a3[:s]
    dn_p, dn_a1, dn_a3 = pos[s:], a1[s:], a3[s:]
    loop_p, loop_a1, loop_a3 = _bridge((up_p[-1], up_a1[-1], up_a3[-1]), (dn_p[0], dn_a1[0], dn_a3[0]), loop_len,
                                       bulge=np.array([0.0, 0.0, 1.0]), back=back)
    P = np.concatenate([up_p, loop_p.reshape(-1, 3), dn_p])
    A1 = np.concatenate([up_a1, loop_a1.reshape(-1, 3), dn_a1])
    A3 = np.concatenate([up_a3, loop_a3.reshape(-1, 3), dn_a3])
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


# ---------------------------------------------------------------- RNA cloverleaf
# Idealised A-form helix for RNA (2.8 A rise = 0.3287 oxDNA length units, 11 bp per
# turn, base-pair planes normal to the axis): backbone sites of consecutive
# nucleotides then sit ~0.65 apart, inside the oxRNA FENE range 0.761 +- 0.25.
RNA_RISE = 0.3287
RNA_TWIST = 2.0 * math.pi / 11.0
RNA_COMPLEMENT = {"A": "U", "U": "A", "G": "C", "C": "G"}


def _rotation_z_to(u, phase):
    """Rotation taking local +z to unit vector u, with a phase rotation about u."""
    u = np.asarray(u, dtype=float)
    u = u / np.linalg.norm(u)
    ref = np.array([0.0, 0.0, 1.0]) if abs(u[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    e1 = np.cross(ref, u)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)
    c, s = math.cos(phase), math.sin(phase)
    x = c * e1 + s * e2
    y = -s * e1 + c * e2
    return np.column_stack([x, y, u])


def _stem_local(seq, rise, twist):
    """Ideal stem along +z from z = 0 (first pair) upwards: strand 0 (seq) up, its
    complement down. Returns pos, a1, a3 of strand 0 and of strand 1."""
    pos, a1, a3, _, _, _, _ = ideal_duplex(seq.replace("U", "T"), rise, twist, COM_RADIUS)
    n = len(seq)
    shift = np.array([0.0, 0.0, 0.5 * (n - 1) * rise])
    return pos[:n] + shift, a1[:n], a3[:n], pos[n:] + shift, a1[n:], a3[n:]


def cloverleaf(seed=0, n_total=76, stems=(4, 7), loops=(3, 8), acceptor=7):
    """One RNA strand of ``n_total`` nt with a seeded random sequence folded into an
    idealised A-form cloverleaf: an acceptor stem closed by the two chain ends and
    three hairpin arms (stems s ~ U{stems} bp, loops ~ U{loops} nt), joined by
    single-stranded linkers that take the remaining nucleotides. The molecule lies in
    the xy plane, arms at -x, +y, +x around the junction, the acceptor stem along -y;
    stem partners are Watson-Crick complements. Returns (System, info) where info
    lists the stems as (first, last, partner_first) index triples and the linker
    lengths; the box is left to the caller (set ``system.box``)."""
    rng = np.random.default_rng(seed)
    s0 = int(acceptor)
    for _ in range(10000):
        arm_s = [int(rng.integers(stems[0], stems[1] + 1)) for _ in range(3)]
        arm_l = [int(rng.integers(loops[0], loops[1] + 1)) for _ in range(3)]
        rest = n_total - 2 * s0 - sum(2 * s + l for s, l in zip(arm_s, arm_l))
        if 8 <= rest <= 16:
            break
    else:
        raise RuntimeError("no cloverleaf layout fits n_total")
    linkers = [rest // 4 + (1 if k < rest % 4 else 0) for k in range(4)]
    alphabet = "ACGU"
    seq_acc = "".join(rng.choice(list(alphabet), size=s0))
    seq_arms = ["".join(rng.choice(list(alphabet), size=s)) for s in arm_s]

    def build(R0, phases, step):
        parts = []  # (pos, a1, a3, bases) in chain order
        # acceptor strand 0 runs toward the junction (local +z -> +y), top at (0, -R0, 0)
        p0, a10, a30, p1, a11, a31 = _stem_local(seq_acc, RNA_RISE, RNA_TWIST)
        Ra = _rotation_z_to([0.0, 1.0, 0.0], phases[0])
        top = np.array([0.0, 0.0, (s0 - 1) * RNA_RISE])
        t = np.array([0.0, -R0, 0.0])
        acc0 = ((p0 - top) @ Ra.T + t, a10 @ Ra.T, a30 @ Ra.T, list(seq_acc))
        acc1 = ((p1 - top) @ Ra.T + t, a11 @ Ra.T, a31 @ Ra.T,
                [RNA_COMPLEMENT[b] for b in reversed(seq_acc)])
        arms = []
        for k, (u, seq, ell) in enumerate(zip(([-1.0, 0, 0], [0, 1.0, 0], [1.0, 0, 0]), seq_arms, arm_l)):
            q0, b10, b30, q1, b11, b31 = _stem_local(seq, RNA_RISE, RNA_TWIST)
            Rk = _rotation_z_to(u, phases[k + 1])
            tk = R0 * np.asarray(u)
            up = (q0 @ Rk.T + tk, b10 @ Rk.T, b30 @ Rk.T)
            dn = (q1 @ Rk.T + tk, b11 @ Rk.T, b31 @ Rk.T)
            lp = _bridge((up[0][-1], up[1][-1], up[2][-1]), (dn[0][0], dn[1][0], dn[2][0]), ell,
                         bulge=np.asarray(u), back=RNA_BACK, step=step)
            arms.append((up, lp, dn, seq))
        parts.append(acc0)
        prev_end = (acc0[0][-1], acc0[1][-1], acc0[2][-1])
        for k, (up, lp, dn, seq) in enumerate(arms):
            lk = _bridge(prev_end, (up[0][0], up[1][0], up[2][0]), linkers[k], bulge=[0.0, 0.0, 1.0],
                         back=RNA_BACK, step=step)
            parts.append((lk[0], lk[1], lk[2], ["U"] * linkers[k]))
            parts.append((up[0], up[1], up[2], list(seq)))
            parts.append((lp[0], lp[1], lp[2], ["U"] * len(lp[0])))
            parts.append((dn[0], dn[1], dn[2], [RNA_COMPLEMENT[b] for b in reversed(seq)]))
            prev_end = (dn[0][-1], dn[1][-1], dn[2][-1])
        lk = _bridge(prev_end, (acc1[0][0], acc1[1][0], acc1[2][0]), linkers[3], bulge=[0.0, 0.0, 1.0],
                     back=RNA_BACK, step=step)
        parts.append((lk[0], lk[1], lk[2], ["U"] * linkers[3]))
        parts.append(acc1)
        pos = np.concatenate([np.asarray(p[0]).reshape(-1, 3) for p in parts])
        a1 = np.concatenate([np.asarray(p[1]).reshape(-1, 3) for p in parts])
        a3 = np.concatenate([np.asarray(p[2]).reshape(-1, 3) for p in parts])
        bases = [b for p in parts for b in p[3]]
        return pos, a1, a3, bases

    def quality(pos, a1, a3):
        # worst deviation of a bonded backbone-site distance from the FENE r0 (oxRNA
        # backbone site -0.4 a1 + 0.2 a3), and the closest approach of backbone sites
        # and of base sites (+0.4 a1) of nucleotides that are not chain neighbours
        back = pos - 0.4 * a1 + 0.2 * a3
        base = pos + 0.4 * a1
        fene = np.abs(np.linalg.norm(np.diff(back, axis=0), axis=1) - 0.761).max()
        n = len(pos)
        idx = np.arange(n)
        far = np.abs(idx[:, None] - idx[None, :]) > 1
        dbb = np.linalg.norm(back[:, None] - back[None], axis=2)[far].min()
        dss = np.linalg.norm(base[:, None] - base[None], axis=2)[far].min()
        return fene, dbb, dss

    best = None
    for step in (0.75, 0.7, 0.8, 0.65):
        for R0 in np.arange(1.6, 4.01, 0.2):
            phases = [0.0, 0.0, 0.0, 0.0]
            for k in range(4):  # greedy phase choice: best worst backbone bond
                cand = []
                for ph in np.linspace(0, 2 * math.pi, 12, endpoint=False):
                    trial = list(phases)
                    trial[k] = ph
                    cand.append((quality(*build(R0, trial, step)[:3])[0], ph))
                phases[k] = min(cand)[1]
            pos, a1, a3, bases = build(R0, phases, step)
            fene, dbb, dss = quality(pos, a1, a3)
            if fene < 0.17 and dbb > 0.6 and dss > 0.3:
                best = (pos, a1, a3, bases)
                break
        if best is not None:
            break
    if best is None:
        raise RuntimeError("could not lay out the cloverleaf without clashes")
    pos, a1, a3, bases = best
    n = len(bases)
    assert n == n_total, (n, n_total)
    quat = io.orientations_to_quaternions(a1, a3, check=False)
    n3 = [i - 1 for i in range(n)]
    n5 = [i + 1 if i < n - 1 else -1 for i in range(n)]
    bonds = [(i, i + 1) for i in range(n - 1)]
    typeid = [io.base_typeid("T" if b == "U" else b) for b in bases]
    system = System(pos - pos.mean(axis=0), quat, typeid, bonds, [4.0 * 30.0] * 3, [0] * n, bases, n3, n5)
    info = dict(arm_stems=arm_s, arm_loops=arm_l, acceptor=s0, linkers=linkers)
    return system, info


def cloverleaf_ensemble(n_replicas, seed=5, box=64.0, d_range=(2.0, 30.0), k_bias=1.0):
    """``n_replicas`` copies of one seeded cloverleaf, each in its own cubic box of
    side ``box``, with a harmonic end-to-end bias k (|r_last - r_first| - d_r)^2 whose
    centres d_r are spaced uniformly over ``d_range``. Returns (systems, ends, k, d0)."""
    base, _ = cloverleaf(seed=seed)
    out = []
    for _ in range(n_replicas):
        out.append(System(base.pos.copy(), base.quat.copy(), base.typeid, base.bonds, [box] * 3, base.strand,
                          base.bases, base.n3, base.n5))
    ends = np.tile([0, base.n - 1], (n_replicas, 1))
    d0 = np.linspace(d_range[0], d_range[1], n_replicas)
    k = np.full(n_replicas, float(k_bias))
    return out, ends, k, d0
