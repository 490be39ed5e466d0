"""Benchmark: oxDNA2 Langevin steps/s on one MI355X, ~8k nucleotides.

Workload (one "step" = one full MD step of the hot path: Langevin half-kick + drift +
NO_SQUISH rotation, neighbour-list skin test / rebuild, every oxDNA2 term in one
fused force evaluation, Langevin second half-kick): 200 ideal 20-bp duplexes with
seeded random sequences, Haar-random orientations and non-overlapping uniform
centres in a cubic box of side 200 (8,000 nt, 1e-3 nt per unit^3), oxDNA2
sequence-averaged, HB + cross-stacking fused, Debye-Hueckel at 0.5 M, T = 296.15 K,
dt = 0.003, gamma = gamma_r = 1. State stays resident in HBM during the timed region.

Prints one JSON line (see the fields below). Options:
    --gpus N     only 1 is supported (one GPU)
    --steps K    timed steps (default 10000)
    --warmup W   untimed steps before timing (default 1000)
    --no-cpu     skip the CPU baseline leg
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hoomd-oxdna_amd"))

FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (published)
TERM_NAMES = ["excvol", "hbond", "cross", "coax", "debye", "hbondcross", "bonded"]


def flops_per_launch(counts, census):
    """Algorithmic flops of one fused force launch: directed evaluation counts per term
    (inside cut-off, nonzero energy) x per-evaluation flops of the op census."""
    total = 0.0
    for t, name in enumerate(TERM_NAMES):
        n_in, n_nz = counts[2 * t], counts[2 * t + 1]
        if n_in == 0:
            continue
        c = census[name]
        c_nz = c["flops_nonzero"] if c["nonzero"] else c["flops_zero"]
        total += (n_in - n_nz) * c["flops_zero"] + n_nz * c_nz
    return total


def traffic_from_profile():
    """HBM bytes per force launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "forces_8k_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def cpu_baseline(system, forces, seconds=12.0):
    """The FP64 CPU oracle on this host: one force evaluation of every term (cell
    list rebuilt each time), single thread, repeated for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.from_system(system, forces)
    orc.compute()
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        orc.compute()
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="steps/s", cores=1, kind="port",
                sample=f"{n} FP64 force evaluations (all oxDNA2 terms, fused HB+CS, cell list rebuilt "
                       f"each time, no integration) of the same {system.n}-nt system, 1 thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the replica / "
              "domain sharding across GPUs is described in DESIGN.md and not launched here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from oxdna import sim, systems
    from oxdna.model import dna1, dna2

    system = systems.duplex_gas(200, 20, box=200.0, seed=1)
    forces, _ = dna2.forces(fused=True)
    s = sim.Simulation(device=sim.GPU(0), seed=1, use_graph=not args.no_graph)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(
        dt=0.003, forces=forces, methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    s.run(max(args.warmup, 1))
    s.synchronize()
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the engine stream sync is what matters
        sync = lambda: None  # noqa: E731
    sync(); s.synchronize()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize(); sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    steps_per_s = args.steps / (t1 - t0)
    e = s.term_energies()
    assert all(np.isfinite(v) for v in e.values()), e

    # dominant kernel: the fused force kernel, timed with HIP events on the engine stream
    import ctypes as C
    lib, ctx = s._lib, s._ctx
    avg_ms = C.c_double()
    assert lib.oxdna_time_forces(ctx, s._mask, 200, C.byref(avg_ms)) == 0
    counts = (C.c_uint64 * 14)()
    assert lib.oxdna_count_evaluations(ctx, s._mask, counts) == 0
    counts = list(counts)
    with open(os.path.join(ROOT, "hoomd-oxdna_amd", "oxdna", "data", "census_flops.json")) as fh:
        census = json.load(fh)["oxdna2_fused"]
    flops = flops_per_launch(counts, census)
    achieved = flops / (avg_ms.value * 1e-3) / 1e12
    pair_evals = sum(counts[2 * t] for t in range(6))
    stats = s.stats()

    line = {
        "metric": "oxDNA2 steps/s @8k nt; batched replica-steps/s vs plain; % FP32 VALU peak",
        "value": round(steps_per_s, 1),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FP64 FENE, min-image, reductions, integrator state)",
        "data": "synthetic",
        "config": {"workload": "oxDNA2 200x20-bp duplexes, 8000 nt, box 200, Langevin dt=0.003, "
                               "fused HB+CS, Debye-Hueckel 0.5 M, 296.15 K",
                   "n_nucleotides": system.n, "graph": not args.no_graph},
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 6),
                     "traffic": traffic_from_profile(), "kernel": "k_forces",
                     "kernel_ms": round(avg_ms.value, 5), "flops_per_launch": flops},
        "pair_evaluations_per_s": round(pair_evals * steps_per_s, 1),
        "nlist_builds": stats["nlist_builds"],
        "energy_per_nt": {k: round(v / system.n, 6) for k, v in e.items()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(system, forces)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
