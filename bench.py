"""Benchmark: oxDNA2 Langevin steps/s on one MI355X, ~8k nucleotides.

Workload (one "step" = one full MD step of the hot path: Langevin half-kick + drift +
NO_SQUISH rotation, neighbour-list skin test / rebuild, every oxDNA2 term in one
fused force evaluation, Langevin second half-kick): 200 ideal 20-bp duplexes with
seeded random sequences, Haar-random orientations and non-overlapping uniform
centres in a cubic box of side 200 (8,000 nt, 1e-3 nt per unit^3), oxDNA2
sequence-averaged, HB + cross-stacking fused, Debye-Hueckel at 0.5 M, T = 296.15 K,
dt = 0.003, gamma = gamma_r = 1. State stays resident in HBM during the timed region.

Then, in the same run ("batched" in the JSON line), two batched replica ensembles in one
persistent launch per 2048 steps, each beside the plain one-system path run replica by
replica: 512 oxDNA2 hairpins (stem 4-30 bp, poly-T loop 4-10 nt, box 30 each) and 256
oxRNA2 76-nt cloverleaves with a harmonic end-to-end bias (centres 2..30, box 64 each).

Prints one JSON line (see the fields below). Options:
    --gpus N           only 1 is supported (one GPU)
    --steps K          timed steps (default 10000)
    --warmup W         untimed steps before timing (default 1000)
    --buffer B         pin the plain run's neighbour-list skin (default: autotuned)
    --no-cpu           skip the CPU baseline leg
    --no-batched       skip the batched ensembles
    --batched-steps S  timed steps of each ensemble (default 2000)
    --plain-sample R   replicas run one by one on the plain path (default 32; 0 skips)
    --copy-probe       only 256 MiB device copies (counter calibration for profiles/)
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hoomd-oxdna_amd"))

FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (published)
HBM_PEAK_BPS = 8.0e12   # MI355X HBM3E (published)
TERM_NAMES = ["excvol", "hbond", "cross", "coax", "debye", "hbondcross", "bonded"]


def flops_per_launch(counts, census, evals_per_pair=1):
    """Algorithmic flops of one force evaluation: evaluation counts per term (inside
    cut-off, nonzero energy) x flops of the op census, which prices one evaluation
    of a pair (forces on both, both torques) or a bond. Counts from the one-system
    kernel are directed (every pair and bond seen from both ends: evals_per_pair=2)
    and are halved, so the figure is the work any evaluation must do, not the
    full-list kernel's duplicated work."""
    total = 0.0
    for t, name in enumerate(TERM_NAMES):
        n_in, n_nz = counts[2 * t] / evals_per_pair, counts[2 * t + 1] / evals_per_pair
        if n_in == 0:
            continue
        c = census[name]
        c_nz = c["flops_nonzero"] if c["nonzero"] else c["flops_zero"]
        total += (n_in - n_nz) * c["flops_zero"] + n_nz * c_nz
    return total


def traffic_from_profile(kernel):
    """HBM bytes per force launch from the committed rocprofv3 counter summary, when it
    was taken on the same kernel instantiation (else None)."""
    path = os.path.join(ROOT, "profiles", "k_forces_fused_oxdna2_8k_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        prof = json.load(fh)
    if not prof.get("kernel", "").startswith(kernel):
        return None
    return prof.get("hbm_bytes_per_launch")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(system, forces, buffer, seconds=8.0):
    """The FP64 CPU oracle on this host, compiled here with g++ -O3 -march=native and
    OpenMP: one evaluation of every term (forces, torques, per-term energies) over a
    neighbour list with the GPU run's skin, built once outside the timed loop (an MD
    code rebuilds it only when a particle has moved half the skin), about `seconds`
    of evaluations at 1 thread and at every core this process may use."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    with tempfile.TemporaryDirectory() as tmp:
        lib, cmd = O.baseline_lib(tmp)
        orc = O.from_system(system, forces)
        r_list = max(float(f.r_cut[("A", "A")]) for f in forces[1:]) + buffer
        pairs = orc.pairs(r_list)
        cores = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
        rates = {}
        for nt in sorted({1, cores}):
            orc.compute_list(pairs, nt, library=lib)
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < seconds:
                orc.compute_list(pairs, nt, library=lib)
                n += 1
            rates[nt] = (n, n / (time.perf_counter() - t0))
    n_all, rate_all = rates[cores]
    return dict(value=round(rate_all, 2), unit="steps/s", cores=cores, kind="port",
                value_1_thread=round(rates[1][1], 2),
                cpu_model=_cpu_model(), nproc=os.cpu_count(),
                compile=cmd,
                sample=f"{n_all} evaluations at {cores} threads and {rates[1][0]} at 1 thread, about {seconds:.0f} s "
                       f"each, of every oxDNA2 term (fused HB+CS, FP64 forces, torques, per-term energies) of the "
                       f"same {system.n}-nt system over a neighbour list of {len(pairs)} pairs (skin {buffer}) built "
                       f"once; no integration (a Langevin update of {system.n} particles is under 1 % of a CPU step)")


def _timed_run(s, steps):
    s.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    s.synchronize()
    return time.perf_counter() - t0


def _langevin(kT):
    from oxdna import sim
    return [sim.Langevin(kT=kT)]


def bench_batched(name, reps, forces, kT, steps, warmup, census, bias=None, sample=32, sample_steps=500):
    """One batched ensemble (all replicas, one persistent launch per chunk of steps)
    against the plain one-system path run replica by replica (a sample of replicas
    spread over the ensemble, each in its own context with its graph-captured step)."""
    import ctypes as C
    import numpy as np
    from oxdna import ensemble, sim
    n_rep = len(reps)
    ens = ensemble.ReplicaEnsemble(device=sim.GPU(0), seed=11)
    ens.create_state_from_snapshots([r.snapshot() for r in reps])
    ens.operations.integrator = sim.Integrator(dt=0.003, forces=forces, methods=_langevin(kT))
    if bias is not None:
        ens.bias = bias
    ens.run(max(warmup, 1))
    dt = _timed_run(ens, steps)
    rate = n_rep * steps / dt
    e = ens.replica_energies()
    assert np.isfinite(e).all(), "non-finite replica energies"
    lib, ctx = ens._lib, ens._ctx
    avg_ms = C.c_double()
    assert lib.oxdna_time_forces(ctx, ens._mask, 200, C.byref(avg_ms)) == 0
    counts = (C.c_uint64 * 14)()
    assert lib.oxdna_count_evaluations(ctx, ens._mask, counts) == 0
    flops = flops_per_launch(list(counts), census)
    achieved = flops / (avg_ms.value * 1e-3) / 1e12
    tuning = ens.tuning()
    # plain path, one replica at a time
    pick = np.linspace(0, n_rep - 1, min(sample, n_rep)).round().astype(int) if sample > 0 else []
    plain_steps, plain_time = 0, 0.0
    for r in pick:
        s = sim.Simulation(device=sim.GPU(0), seed=11 + int(r))
        s.create_state_from_snapshot(reps[r].snapshot())
        s.operations.integrator = sim.Integrator(dt=0.003, forces=forces, methods=_langevin(kT))
        s.run(64)
        plain_time += _timed_run(s, sample_steps)
        plain_steps += sample_steps
        del s
    plain_rate = plain_steps / plain_time if plain_steps else None
    sizes = np.diff(ens.offsets)
    return {
        "workload": name,
        "replicas": n_rep,
        "nucleotides": int(sizes.sum()),
        "replica_size_min_max": [int(sizes.min()), int(sizes.max())],
        "steps": steps,
        "replica_steps_per_s": round(rate, 1),
        "ensemble_steps_per_s": round(steps / dt, 1),
        "plain_per_replica_replica_steps_per_s": round(plain_rate, 1) if plain_rate else None,
        "plain_sample": f"{len(pick)} replicas spread over the ensemble, {sample_steps} steps each, "
                        "one context per replica (graph-captured step), run one after another",
        "speedup_vs_plain": round(rate / plain_rate, 2) if plain_rate else None,
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 6), "traffic": None,
                     "kernel": f"k_ensemble<{tuning['replica_workgroup']}, false, false> (forces only)",
                     "kernel_ms": round(avg_ms.value, 5), "flops_per_launch": flops},
        "tuning": {"replica_workgroup": tuning["replica_workgroup"],
                   "replica_records_per_pass": tuning["replica_records_per_pass"]},
        "energy_per_nt": {k: round(float(v) / int(sizes.sum()), 6)
                          for k, v in zip(ensemble.REPLICA_ENERGY_NAMES, e.sum(axis=0))},
    }


def copy_probe(mib=256, reps=5):
    """Device-to-device copies of a known size: the reference reading for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run (bench.py --copy-probe)."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": mib << 20, "copies": reps}), flush=True)


def bench_large(model_name, steps, warmup, census_all):
    """4096 ideal 64-bp duplexes (524,288 nt) in a box of 640 (2e-3 nt per unit^3),
    cell list with a 10 % skin, fused Langevin steps."""
    import ctypes as C
    from oxdna import sim, systems
    from oxdna.model import dna1, dna2
    model = {"dna1": dna1, "dna2": dna2}[model_name]
    t0 = time.perf_counter()
    system = systems.duplex_gas(4096, 64, box=640.0, seed=3)
    t_gen = time.perf_counter() - t0
    forces, nl = model.forces(fused=True)
    nl.buffer = 0.1 * max(float(f.r_cut[("A", "A")]) for f in forces[1:])
    s = sim.Simulation(device=sim.GPU(0), seed=1)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(
        dt=0.003, forces=forces, methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    s.run(max(warmup, 1))
    dt = _timed_run(s, steps)
    lib, ctx = s._lib, s._ctx
    avg_ms = C.c_double()
    assert lib.oxdna_time_forces(ctx, s._mask, 20, C.byref(avg_ms)) == 0
    counts = (C.c_uint64 * 14)()
    assert lib.oxdna_count_evaluations(ctx, s._mask, counts) == 0
    census = census_all["oxdna1_fused" if model_name == "dna1" else "oxdna2_fused"]
    flops = flops_per_launch(list(counts), census, evals_per_pair=2)
    achieved = flops / (avg_ms.value * 1e-3) / 1e12
    tpp = C.c_int()
    assert lib.oxdna_get_tuning(ctx, C.byref(tpp), None) == 0
    st = s.stats()
    return {
        "workload": f"{model_name}: 4096 x 64-bp duplexes, 524288 nt, box 640, skin {nl.buffer:.3f}, "
                    "Langevin dt=0.003, fused",
        "steps": steps, "steps_per_s": round(steps / dt, 1), "ms_per_step": round(dt * 1e3 / steps, 4),
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 6), "traffic": None,
                     "kernel": f"k_forces<{256 // tpp.value}>", "kernel_ms": round(avg_ms.value, 4),
                     "flops_per_launch": flops},
        "pair_evaluations_per_s": round(sum(counts[2 * t] for t in range(6)) / 2 * steps / dt, 1),
        "nlist_builds": st["nlist_builds"], "max_neigh": st["max_neigh"], "generate_s": round(t_gen, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--buffer", type=float, default=0.0,
                    help="neighbour-list skin of the plain run (0: autotune over 0.1 ... 0.8)")
    ap.add_argument("--lanes", type=int, default=0,
                    help="force-kernel lanes per particle (4, 8, 16, 32) for the plain run; 0: autotune")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched replica workloads")
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="replicas run one by one on the plain path for the comparison (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 256 MiB device copies (counter calibration for profiles)")
    ap.add_argument("--large", action="store_true",
                    help="only the 524,288-nt oxDNA2 and oxDNA1 systems (one JSON line)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.large:
        with open(os.path.join(ROOT, "hoomd-oxdna_amd", "oxdna", "data", "census_flops.json")) as fh:
            census_all = json.load(fh)
        runs = [bench_large(m, args.steps, args.warmup, census_all) for m in ("dna2", "dna1")]
        print(json.dumps({"metric": "oxDNA steps/s @524k nt", "value": runs[0]["steps_per_s"],
                          "unit": "steps/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "fp32 (FP64 FENE, min-image, reductions, integrator state)",
                          "data": "synthetic", "runs": runs}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the replica / "
              "domain sharding across GPUs is not launched here.",
              file=sys.stderr)
        sys.exit(2)

    # torch (plumbing only) is initialised first: its set-up takes seconds of host time,
    # which must not fall between the warm-up steps and the timed ones
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
        sync()
    except Exception:  # the engine stream sync is what matters
        sync = lambda: None  # noqa: E731
    import numpy as np
    from oxdna import sim, systems
    from oxdna.model import dna1, dna2

    system = systems.duplex_gas(200, 20, box=200.0, seed=1)
    forces, nl = dna2.forces(fused=True)
    if args.buffer > 0:
        nl.buffer = args.buffer
    s = sim.Simulation(device=sim.GPU(0), seed=1, use_graph=not args.no_graph, lanes_per_particle=args.lanes)
    s.create_state_from_snapshot(system.snapshot())
    s.operations.integrator = sim.Integrator(
        dt=0.003, forces=forces, methods=[sim.Langevin(kT=dna1.kT_from_temperature(296.15))])
    # the neighbour-list skin from the autotuner (buffers 0.1 ... 0.8, 400 untimed and
    # 2000 timed steps each from the initial state, which it restores), unless --buffer
    # pins it; then the warm-up steps, right before the timed ones
    s.run(0)
    skin_rates = None if args.buffer > 0 else s.tune_nlist_buffer(steps=2000)
    s.run(max(args.warmup, 1))
    sync(); s.synchronize()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize(); sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    steps_per_s = args.steps / (t1 - t0)
    e = s.term_energies()
    assert all(np.isfinite(v) for v in e.values()), e

    # dominant kernel: the run step's own launch (k_forces FUSED: every term + the
    # Langevin integrator in its epilogue), timed with HIP events on the engine stream
    # around graph replays of run steps; beside it the forces-only instantiation
    import ctypes as C
    lib, ctx = s._lib, s._ctx
    counts = (C.c_uint64 * 14)()
    assert lib.oxdna_count_evaluations(ctx, s._mask, counts) == 0
    counts = list(counts)
    with open(os.path.join(ROOT, "hoomd-oxdna_amd", "oxdna", "data", "census_flops.json")) as fh:
        census = json.load(fh)["oxdna2_fused"]
    flops = flops_per_launch(counts, census, evals_per_pair=2)
    step_ms = C.c_double()
    assert lib.oxdna_time_steps(ctx, s._mask, 1 + 4 * 96, C.byref(step_ms)) == 0
    fo_ms = C.c_double()
    assert lib.oxdna_time_forces(ctx, s._mask, 200, C.byref(fo_ms)) == 0
    achieved = flops / (step_ms.value * 1e-3) / 1e12
    achieved_fo = flops / (fo_ms.value * 1e-3) / 1e12
    pair_evals = sum(counts[2 * t] for t in range(6)) // 2
    tpp = C.c_int()
    assert lib.oxdna_get_tuning(ctx, C.byref(tpp), None) == 0
    P = 256 // tpp.value
    kname = f"k_forces<{P},0,0,0,FUSED,FAR2>"
    traffic = traffic_from_profile(kname)
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
        "tuning": dict(s.tuning(), nlist_buffer_steps_per_s=(
            {str(k): round(v, 1) for k, v in skin_rates.items()} if skin_rates else "pinned by --buffer")),
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 6),
                     "traffic": traffic,
                     "hbm_frac": (round(traffic / (step_ms.value * 1e-3) / HBM_PEAK_BPS, 5)
                                  if traffic else None),
                     "kernel": kname + " (the run step's launch: forces + Langevin integrator)",
                     "kernel_ms": round(step_ms.value, 5), "flops_per_launch": flops},
        "roofline_forces_only": {"kernel": f"k_forces<{P},0,0,0,0,FAR2> (forces only, oxdna_time_forces)",
                                 "kernel_ms": round(fo_ms.value, 5), "achieved": round(achieved_fo, 4),
                                 "frac": round(achieved_fo / FP32_PEAK_TFLOPS, 6)},
        "pair_evaluations_per_s": round(pair_evals * steps_per_s, 1),
        "nlist_builds": stats["nlist_builds"],
        "energy_per_nt": {k: round(v / system.n, 6) for k, v in e.items()},
    }
    if not args.no_batched:
        from oxdna import ensemble
        from oxdna.model import rna2
        kT = dna1.kT_from_temperature(296.15)
        hp = systems.hairpin_ensemble(512, seed=3, box=30.0)
        fh, _ = dna2.forces(fused=True)
        b1 = bench_batched("512 oxDNA2 hairpins (stem 4-30 bp, poly-T loop 4-10 nt), box 30 each, "
                           "Langevin dt=0.003, fused HB+CS, 0.5 M, 296.15 K", hp, fh, kT,
                           args.batched_steps, 200, census, sample=args.plain_sample)
        cl, ends, kb, d0 = systems.cloverleaf_ensemble(256, seed=5, box=64.0, d_range=(2.0, 30.0), k_bias=0.02)
        fr, _ = rna2.forces(fused=True)
        with open(os.path.join(ROOT, "hoomd-oxdna_amd", "oxdna", "data", "census_flops.json")) as fh_:
            census_rna = json.load(fh_).get("oxrna2_fused", census)
        b2 = bench_batched("256 oxRNA2 76-nt cloverleaves, harmonic end-to-end bias k=0.02, centres 2..30, "
                           "box 64 each, Langevin dt=0.003, 1 M, 296.15 K", cl, fr, kT, args.batched_steps, 200,
                           census_rna, bias=ensemble.HarmonicBias(ends, kb, d0), sample=args.plain_sample)
        line["batched"] = [b1, b2]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(system, forces, buffer=nl.buffer)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
