"""Benchmark of the spin-dynamics hot path on one MI355X.

Workload (default): the batched ALC field sweep -- muon + electron + 4 1H +
14N (Hilbert dimension 192), 200 longitudinal fields in 0-3 T x 500 powder
orientations = 1e5 Hermitian 192 x 192 eigensolves, integrate_decaying with
tau = 2.197 us, synthetic couplings from a fixed seed.  One step = the whole
sweep (H build -> eigh -> basis change -> integral -> powder accumulation)
with inputs resident in HBM.

Prints one JSON line: orientation-solves/s (whole job), per-kernel roofline
of the dominant kernel from HIP events recorded in the timed region, and the
FP64 oracle timed on the host cores over a stride sample of the same sweep
(which also gives max|dP| of the GPU result on that sample).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3      # dense FP32 matrix (and vector) peak of MI355X
HBM_PEAK_GBS = 8000.0
# kernels whose algorithmic work is priced against the FP32 peak
FP32_KERNELS = ("k_timesum_mfma", "k_tridiag32")


def algorithmic(kernel, n, ntimes, real_mode=True):
    """(unit, amount per instance) for the roofline: real flops per instance
    (complex multiply-add = 8), or HBM bytes for pure data-movement kernels."""
    if kernel == "k_tridiag":
        return "flop", 16.0 / 3.0 * n ** 3
    if kernel == "k_backtransform":
        return "flop", 8.0 * n ** 3
    if kernel == "k_gemm_cn":
        return "flop", 8.0 * n ** 3
    if kernel == "k_eigh_jacobi":
        return "flop", (16.0 / 3.0 + 8.0) * n ** 3
    if kernel == "k_lindblad_modes":
        # Liouvillian order N = n^2: Hessenberg + Schur + eigenvectors (LAPACK's
        # zgeev count, ~100 N^3 real flops) and the mode amplitudes (8/3 N^3)
        return "flop", (100.0 + 8.0 / 3.0) * float(n * n) ** 3
    if kernel == "k_stedc":
        # the merge products Q_blockdiag Z: n^3 at the top level, n^3/4 below, ...
        return "flop", 4.0 / 3.0 * n ** 3
    if kernel == "k_timesum":
        return "flop", (4.0 if real_mode else 8.0) * n * n * ntimes
    if kernel in ("k_timesum_mfma", "k_timesum_mfma64"):
        # Y = C [cos | sin]: n x n x 2T real multiply-adds (FP32 / FP64)
        return "flop", 4.0 * n * n * ntimes
    if kernel == "k_tridiag32":
        return "flop", 16.0 / 3.0 * n ** 3
    if kernel == "k_phase_table":
        # cos and sin tables, FP32, written once
        return "byte", 8.0 * n * ntimes
    if kernel == "k_integral":
        return "flop", 20.0 * n * n
    if kernel == "k_build_h":
        return "byte", 16.0 * n * n
    if kernel == "k_apply_muon":
        return "byte", 2 * 16.0 * n * n
    if kernel == "k_form_c":
        return "byte", 24.0 * n * n
    return "byte", 0.0


def measured_traffic(tag, kernel):
    """HBM bytes per instance of `kernel` from the committed counter summary
    (FETCH_SIZE / WRITE_SIZE scaled by a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{tag}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    for name, rec in data["kernels"].items():
        if name == kernel or name.startswith(kernel + "_") or name.startswith(kernel + "<"):
            return rec["fetch_bytes_per_instance"] + rec["write_bytes_per_instance"]
    return None


def workload(name):
    from muspinsim_amd import workloads as W
    if name == "c2":
        return W.c2_alc_sweep(), "alc_sweep_n192_200fields_x_500orient_integral", "alc_sweep_n192"
    if name == "c1":
        return W.c1_zero_field_powder(), "zero_field_powder_n16_100orient_1000t", "zero_field_n16"
    if name == "c3":
        return W.c3_tf_stress(), "tf_stress_n512_256orient_4000t", "tf_stress_n512"
    if name == "c5":
        return W.c5_fit(), "fit_forward_n64_256orient_2000t", "fit_forward_n64"
    if name == "c4":
        return W.c4_uneven_batch(), "uneven_batch_n4_to_1024_4096inst_101t", "uneven_batch"
    if name == "c4d":
        return W.c4_dissipative(), "lindblad_n8_N64_1024inst_2000t", "lindblad_n8"
    raise SystemExit(f"unknown workload {name}")


def time_plan(eng, plan, args, profile=True):
    """W untimed steps, then K timed steps bracketed by synchronisations;
    per-kernel HIP-event times of the timed steps (graph replays record no
    events, so with --graph one extra eager step supplies them), and one more
    untimed step without event bracketing for the overhead-free rate."""
    import torch
    from muspinsim_amd import _native as N
    run = eng.execute_graph if args.graph else eng.execute
    for _ in range(args.warmup):
        run(plan)
    torch.cuda.synchronize()
    N.profile_report()                      # clear
    N.load_library().msp_profile_enable(1 if (profile and not args.graph) else 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(plan)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    results = plan.results.cpu().numpy()
    rows = plan.rows.cpu().numpy()
    prof_steps = args.steps
    if profile and args.graph:
        N.load_library().msp_profile_enable(1)
        eng.execute(plan)
        torch.cuda.synchronize()
        prof_steps = 1
    N.load_library().msp_profile_enable(0)
    prof = N.profile_report()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run(plan)
    torch.cuda.synchronize()
    dt_clean = time.perf_counter() - t1
    return dt, results, rows, prof, prof_steps, dt_clean


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-sample", type=int, default=24000,
                    help="instances of the sweep timed on the host (0 = skip)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="0 = the job's CPU share (affinity set, limited by OMP_NUM_THREADS)")
    ap.add_argument("--chunk", type=int, default=0, help="instances per launch (0 = auto)")
    ap.add_argument("--chunk-gib", type=float, default=0.0,
                    help="device memory budget of one chunk in GiB (0 = engine default)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="mixed: time sums of the real routes on FP32 matrix cores; the plain "
                         "FP64 path is timed beside it for the speedup")
    ap.add_argument("--no-plain", action="store_true",
                    help="with --precision mixed, skip the FP64 comparison run (profiling)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the plan from a captured HIP graph (SpinEngine.execute_graph)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy a device buffer of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (instances "
              "are independent and would shard by contiguous ranges, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "celio":
        return celio_bench(args)

    import numpy as np
    (systems, batch), wname, tag = workload(args.workload)
    dims = np.array([systems[int(k)].n for k in batch.sys_id])
    n = int(dims.max())
    ninst = len(batch)
    width = 1 if batch.y == "integral" else len(batch.times)

    # ---- CPU baseline first: forked workers, before anything touches the GPU
    cpu = None
    cpu_rows = None
    sample = None
    if args.cpu_sample > 0:
        from oracle import baseline
        stride = max(1, ninst // args.cpu_sample)
        sample = np.arange(0, ninst, stride)[: args.cpu_sample]
        procs = min(args.cpu_procs or baseline.cpu_share(), len(sample))
        cpu_rows, wall, busy = baseline.run(systems, batch, sample, procs)
        one = sample[: max(1, min(len(sample), 64))]
        _, wall1, _ = baseline.run(systems, batch, one, 1)
        host = baseline.host_info()
        rate = len(sample) / wall
        cpu = {
            "value": rate,
            "unit": "orientation-solves/s",
            "cores": procs,
            "kind": "port",
            "sample": f"{len(sample)} of {ninst} instances (stride {stride}) of the same sweep, "
                      f"numpy FP64 eigh + basis change + integral, {procs} processes x 1 BLAS "
                      f"thread (instances[r::P], summed) on {host['model']}",
            "single_core_value": len(one) / wall1,
            "host": host,
            # the reference's MPI split scales at best linearly in processes:
            # the measured per-process rate times every CPU of the affinity
            # set bounds what the whole host could reach
            "full_host_linear_bound": rate / procs * host["affinity_cpus"],
        }

    # ---- GPU
    import torch
    from muspinsim_amd import _native as N
    from muspinsim_amd.engine import SpinEngine

    kw = {}
    if args.chunk:
        kw["max_chunk"] = args.chunk
    if args.chunk_gib:
        kw["chunk_bytes"] = int(args.chunk_gib * (1 << 30))
    eng = SpinEngine("cuda:0", precision=args.precision, **kw)
    plan = eng.prepare(systems, batch)
    dt, results, rows, prof, prof_steps, dt_clean = time_plan(eng, plan, args)
    plain = None
    if args.precision != "fp64" and not args.no_plain:
        # the plain FP64 path on the same workload, same steps, for the speedup
        del plan
        torch.cuda.empty_cache()
        eng64 = SpinEngine("cuda:0", precision="fp64", **kw)
        plan64 = eng64.prepare(systems, batch)
        dt64, res64, rows64, _, _, _ = time_plan(eng64, plan64, args, profile=False)
        plain = {"value": ninst * args.steps / dt64, "ms_per_step": dt64 / args.steps * 1e3,
                 "max_abs_dP_mixed_vs_fp64_gpu_all_instances": float(np.abs(rows - rows64).max())}
        del plan64
        torch.cuda.empty_cache()

    value = ninst * args.steps / dt
    # dominant kernel and its roofline
    kname = max(prof, key=lambda k: prof[k][1])
    launches, tot_ms = prof[kname]
    # algorithmic work summed over the instances (their dimensions may differ)
    unit = algorithmic(kname, n, width)[0]
    work = sum(algorithmic(kname, int(d), width)[1] * c
               for d, c in zip(*np.unique(dims, return_counts=True)))
    per_inst = work / ninst
    inst_per_launch = ninst * prof_steps / launches
    avg_s = tot_ms / launches / 1e3
    if unit == "flop":
        achieved = per_inst * inst_per_launch / avg_s / 1e12
        peak = FP32_PEAK_TFLOPS if kname in FP32_KERNELS else FP64_PEAK_TFLOPS
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None}
    else:
        achieved = per_inst * inst_per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    per_inst_bytes = measured_traffic(tag, kname)
    if per_inst_bytes is not None:
        roof["traffic"] = per_inst_bytes * inst_per_launch
        roof["traffic_source"] = f"profiles/{tag}_hbm_traffic.json"
    roof["kernel"] = kname
    roof["avg_launch_ms"] = tot_ms / launches
    roof["instances_per_launch"] = inst_per_launch
    maxdp = None
    if cpu_rows is not None:
        maxdp = float(np.abs(rows[sample] - cpu_rows).max())
    line = {
        "metric": "orientation-solves/s",
        "value": value,
        "unit": "orientation-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32-mfma",
        "data": "synthetic (seeded couplings and orientations)",
        "config": {"workload": wname, "launch": "hip-graph replay" if args.graph else "stream",
                   "tuning": "muspinsim_amd/tuning.json" if eng.tuning else "built-in", "hilbert_dim": n if dims.min() == n else
                   f"{int(dims.min())}-{n}", "instances": ninst,
                   "precision": ("fp64-plain" if args.precision == "fp64" else
                                 "mixed: FP64 eigensolver and basis change, time sum on FP32 "
                                 "MFMA with FP64 phase reduction and FP64 row sums"),
                   "eigensolver": "hessenberg+complex-schur (Liouvillian)"
                   if any(getattr(sy, "dissip", None) for sy in systems)
                   else "householder+divide-and-conquer" if dims.min() > 64
                   else "jacobi-lds" if n <= 64 else "jacobi-lds + householder+divide-and-conquer"},
        "max_abs_dP_vs_fp64_oracle": maxdp,
        "value_without_event_timing": ninst / dt_clean,
        "kernels_ms_per_step": {k: v[1] / prof_steps for k, v in sorted(prof.items())},
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "gpu_vs_cpu_full_host_bound": (value / cpu["full_host_linear_bound"]) if cpu else None,
        "fp64_plain_gpu": plain,
        "speedup_vs_fp64_plain_gpu": (value / plain["value"]) if plain else None,
        "results_checksum": float(np.sum(results)),
    }
    if args.workload == "c5":
        line["adjoint"] = adjoint_timing(systems, batch, args.steps)
    print(json.dumps(line), flush=True)


def celio_bench(args):
    """Celio/Trotter workload: mu + 4 51V (n = 8192), 377 ZCW orientations x
    4 random-phase states, k = 10, 50 times; one step = one msp_celio_evolve
    over all 1508 states with the factors, tables and phases resident in HBM.
    The value counts orientations (each a full 50-point polarisation series
    averaged over its 4 states).  CPU baseline: the oracle (numpy, one core)
    on 2 of the states with the same phases, which also gives max|dP|."""
    import numpy as np
    from muspinsim_amd import celio, workloads as W
    from muspinsim_amd.experiment import ExperimentRunner
    celio.seed(0)
    cfg = W.celio_mu4v_config()
    er = ExperimentRunner(cfg)
    snaps = cfg[0:len(cfg)]
    job, combine, nt = er._celio_job(snaps)
    nori = len(snaps)
    cpu = None
    ref = {}
    if args.cpu_sample > 0:
        from oracle import celio as oc
        from oracle import physics
        sample = [0, job.nstates // 2 + 1][: max(1, min(2, args.cpu_sample))]
        t = np.asarray(snaps[0].t, dtype=float)
        t0 = time.perf_counter()
        for sidx in sample:
            snap = snaps[job._state_cfg[sidx]]
            er.load_config(snap)
            terms = []
            for tm in er.system.terms:
                idx = [int(i) for i in np.atleast_1d(tm.indices)]
                T = np.asarray(tm.tensor, dtype=float)
                terms.append(("linear", idx[0], idx[0], T) if len(idx) == 1 else
                             ("bilinear", idx[0], idx[1], T))
            terms += [("linear", i, i, er.B * g) for i, g in enumerate(er.system.gammas)]
            ref[sidx] = oc.fast_evolve(er.system.Is, terms, cfg.celio_k, t,
                                       physics.sigma_mu(er.p), job._phases[sidx][None])
        wall = time.perf_counter() - t0
        per_state = wall / len(sample)
        cpu = {"value": 1.0 / (per_state * cfg.celio_averages), "unit": "orientation-solves/s",
               "cores": 1, "kind": "port",
               "sample": f"{len(sample)} of {job.nstates} states (orientations {sorted(set(job._state_cfg[i] for i in sample))}), "
                         "oracle numpy Trotter loop, one core; value = 1 / (time per state x "
                         f"{cfg.celio_averages} states per orientation)"}

    import torch
    from muspinsim_amd import _native as N
    prep = job.prepare()
    out = torch.empty((job.nstates, nt), dtype=torch.float64, device=prep["eng"].device)
    for _ in range(args.warmup):
        job.launch(prep, nt, out)
    torch.cuda.synchronize()
    N.profile_report()
    N.load_library().msp_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.launch(prep, nt, out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    N.load_library().msp_profile_enable(0)
    prof = N.profile_report()
    launches, tot_ms = prof["k_celio_evolve"]
    res = out.cpu().numpy()
    flops = job.flops_per_state(nt) * job.nstates
    achieved = flops / (tot_ms / launches / 1e3) / 1e12
    maxdp = None
    if ref:
        maxdp = float(max(np.abs(0.5 * res[s] - r).max() for s, r in ref.items()))
    rows, s0 = [], 0
    for wts, w in combine:
        rows.append((wts @ res[s0:s0 + len(wts)]) * w)
        s0 += len(wts)
    value = nori * args.steps / dt
    line = {
        "metric": "orientation-solves/s", "value": value, "unit": "orientation-solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (fixed couplings, seeded phases)",
        "config": {"workload": "celio_mu4V_n8192_377orient_x4states_k10_50t", "hilbert_dim": job.n,
                   "instances": nori, "states": job.nstates, "trotter_k": cfg.celio_k,
                   "factors_per_trotter_step": len(job._desc) // len(job._cfg_steps),
                   "launch": "stream", "precision": "fp64-plain",
                   "method": "celio random-state (state in LDS)"},
        "max_abs_dP_vs_fp64_oracle": maxdp,
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in sorted(prof.items())},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_celio_evolve", "avg_launch_ms": tot_ms / launches,
                     "instances_per_launch": nori,
                     "note": "FP64 vector FMA work, 8 n m flops per factor application"},
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "results_checksum": float(np.sum(rows)),
    }
    per_inst_bytes = measured_traffic("celio_mu4V_n8192", "k_celio_evolve")
    if per_inst_bytes is not None:
        line["roofline"]["traffic"] = per_inst_bytes * nori
        line["roofline"]["traffic_source"] = "profiles/celio_mu4V_n8192_hbm_traffic.json"
    print(json.dumps(line), flush=True)


def adjoint_timing(systems, batch, steps):
    """Cost of one chi^2 gradient of the fit configuration (10 parameters:
    6 hyperfine components, 4 muon-proton distances) by the adjoint pass, next
    to the forward pass it is built on.  The reference's L-BFGS-B takes
    finite differences, n_params + 1 forward simulations per gradient
    (muspinsim/fitting.py:99-105)."""
    import numpy as np
    import torch
    from muspinsim_amd import workloads as W
    from muspinsim_amd.adjoint import PolarisationModel
    _, p0 = W.c5_system(None)
    model = PolarisationModel(lambda th: W.c5_system(th)[0], W.c5_dterms, batch)
    target = model.forward(p0).cpu().numpy()
    r = np.zeros_like(target) + 1e-3
    model.vjp(p0, r)                        # warm-up (plans, buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        model.forward(p0)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        model.vjp(p0, r)
    torch.cuda.synchronize()
    vjp = (time.perf_counter() - t0) / steps
    nparam = len(p0)
    return {"params": nparam, "forward_ms": fwd * 1e3, "gradient_ms": vjp * 1e3,
            "finite_difference_ms": (nparam + 1) * fwd * 1e3,
            "gradient_speedup_vs_finite_differences": (nparam + 1) * fwd / vjp}


if __name__ == "__main__":
    main()
