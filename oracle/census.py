"""Op census of the oxDNA terms (TEST / MEASUREMENT INFRASTRUCTURE).

Runs the oracle's restatement on a counting scalar over a workload and writes the
mean flops (and special functions: acos / sin / exp / log) per evaluation, split by
whether the evaluation's energy is zero (the reference exits early there) or not,
for every term. The engine's measurement path multiplies these per-evaluation
figures by the evaluation counts its own kernel reports.

    python oracle/census.py   # writes hoomd-oxdna_amd/oxdna/data/census_flops.json
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [os.path.join(ROOT, "hoomd-oxdna_amd"), HERE]

import oracle as O  # noqa: E402
from oxdna import systems  # noqa: E402
from oxdna.model import dna1, dna2, rna2  # noqa: E402

NAMES = ["excvol", "hbond", "cross", "coax", "debye", "hbondcross", "bonded"]


def census(systems_, forces):
    c = sum(O.from_system(s, forces).census() for s in systems_)
    out = {}
    for t, name in enumerate(NAMES):
        n_in, n_nz, f0, f1, s0, s1 = c[t]
        if n_in == 0:
            continue
        nz0 = n_in - n_nz
        out[name] = dict(
            evaluations=int(n_in), nonzero=int(n_nz),
            flops_zero=(f0 / nz0) if nz0 else 0.0, flops_nonzero=(f1 / n_nz) if n_nz else 0.0,
            special_zero=(s0 / nz0) if nz0 else 0.0, special_nonzero=(s1 / n_nz) if n_nz else 0.0)
    return out


def nicked_helix(n_bp=20, copies=4):
    """Ideal duplexes stacked end to end into one continuous helix with nicks, so that
    the nicked neighbours coaxially stack."""
    import numpy as np
    from oxdna import io
    rng = np.random.default_rng(9)
    P, A1, A3, T, B = [], [], [], [], []
    off = 0
    for k in range(copies):
        pos, a1, a3, bases, _, _, n5 = systems.ideal_duplex(systems.random_sequence(rng, n_bp))
        ang = k * n_bp * systems.TWIST
        R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
        shift = np.array([0, 0, k * n_bp * systems.RISE])
        P.append(pos @ R.T + shift); A1.append(a1 @ R.T); A3.append(a3 @ R.T)
        T += [io.base_typeid(b) for b in bases]
        B += [(i + off, n5[i] + off) for i in range(len(bases)) if n5[i] >= 0]
        off += len(bases)
    P = np.concatenate(P)
    P -= P.mean(axis=0)
    q = io.orientations_to_quaternions(np.concatenate(A1), np.concatenate(A3))
    n = len(T)
    return systems.System(P, q, T, B, [60.0] * 3, [0] * n, ["A"] * n, [-1] * n, [-1] * n)


def main():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers as H
    # thermal-like perturbations of the benchmark gas, plus a dense gas so that
    # contacts (nonzero excluded volume, coaxial stacking) are sampled too
    work = [H.jiggle(systems.duplex_gas(200, 20, box=200.0, seed=1), seed=2, dx=0.05, dq=0.08),
            H.jiggle(systems.duplex_gas(60, 20, box=28.0, seed=3, min_dist=0.9), seed=4, dx=0.05, dq=0.08)]
    work.append(H.jiggle(nicked_helix(), seed=5, dx=0.02, dq=0.03))
    res = {}
    for label, model in (("oxdna2", dna2), ("oxdna1", dna1), ("oxrna2", rna2)):
        res[label + "_unfused"] = census(work, model.forces(fused=False)[0])
        res[label + "_fused"] = census(work, model.forces(fused=True)[0])
    res["_workload"] = ("200 20-bp duplexes (8000 nt) in a box of 200 and 60 20-bp duplexes in a box of 28, "
                        "both perturbed from ideal (0.05 position, 0.08 rad rotation noise)")
    res["_rule"] = ("add/sub/mul/div/sqrt = 1 flop (FMA = 2), acos/sin/exp/log = 1 special op; "
                    "per unordered-pair (per bond) evaluation by the FP64 restatement")
    out = os.path.join(ROOT, "hoomd-oxdna_amd", "oxdna", "data", "census_flops.json")
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps(res["oxdna2_fused"], indent=1))


if __name__ == "__main__":
    main()
