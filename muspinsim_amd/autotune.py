"""Autotuner for the engine's launch choices on the device at hand.

The reference makes one choice for every size (numpy.linalg.eigh per
instance, muspinsim/spinop.py:69, inside the loop of experiment.py:369).
Here the batch runs through different kernels depending on the Hilbert
dimension n and the number of instances of that dimension, and the best
choice is measured, not assumed:

  eigh          for each bucket n (and a ladder of batch sizes), the LDS
                Jacobi eigensolver (n <= 64) or Householder + divide and
                conquer: Jacobi holds one matrix per workgroup for the whole
                solve, so it wins on latency for small batches and loses on
                throughput for large ones.
  timesum       the smallest n from which the mixed-precision time sum
                (FP32 matrix cores) beats the FP64 one on the same batch, and
                the smallest n from which the FP64 matrix-core time sum beats
                the FP64 VALU one.
  max_chunk     instances per launch of the n = 192 divide-and-conquer path.

`tune()` times them with HIP events around the library calls and returns a
dict; `main()` writes it to JSON (muspinsim_amd/tuning.json by default),
which SpinEngine(tuning=...) reads at plan time.  Usage on the GPU:

    python -m muspinsim_amd.autotune [--out PATH] [--quick]
"""

import argparse
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")

EIGH_NS = (4, 8, 12, 16, 24, 32, 48, 64)
EIGH_BATCHES = (64, 256, 1024, 4096, 16384)
TIMESUM_NS = (16, 24, 32, 48, 64, 96, 128)


def _events_ms(fn, reps):
    import torch
    fn()                                   # warm up (kernel attributes, caches)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def _herm(n, batch, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.randn((batch, n, n), dtype=torch.complex128, device="cuda", generator=g)
    return (X + X.conj().transpose(1, 2).contiguous()) / 2


def time_eigh(lib, n, batch, method, reps=3):
    """Best-of-reps milliseconds of one msp_eigh call on random Hermitian
    matrices (the input is restored from a copy before every call)."""
    import torch
    from . import _native as N
    H0 = _herm(n, batch, 1000 + n)
    H = torch.empty_like(H0)
    E = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    U = torch.empty_like(H0)
    info = torch.empty((batch,), dtype=torch.int32, device="cuda")
    wsb = lib.msp_eigh_workspace(n, batch, method)
    work = torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def call():
        H.copy_(H0)
        N.check(lib.msp_eigh(n, batch, H.data_ptr(), E.data_ptr(), U.data_ptr(), info.data_ptr(),
                             method, work.data_ptr(), wsb, st))

    copy_ms = _events_ms(lambda: H.copy_(H0), reps)
    return max(_events_ms(call, reps) - copy_ms, 1e-4)


def eigh_ladders(raw, tol=0.05):
    """{n: [[batch_min, method], ...]} from timings {"<n>x<batch>": {"jacobi_ms",
    "dc_ms"}}: the faster eigensolver from each batch size on, switching only
    when the other method is faster by more than `tol` (differences within the
    run-to-run noise keep the previous choice)."""
    by_n = {}
    for key, t in raw.items():
        n, batch = (int(x) for x in key.split("x"))
        by_n.setdefault(n, []).append((batch, t["jacobi_ms"], t["dc_ms"]))
    table = {}
    for n, rows in sorted(by_n.items()):
        ladder = []
        for batch, tj, td in sorted(rows):
            if not ladder:
                ladder.append([1, "jacobi" if tj <= td else "dc"])
                continue
            cur = ladder[-1][1]
            t_cur, t_other = (tj, td) if cur == "jacobi" else (td, tj)
            if t_other < t_cur * (1.0 - tol):
                ladder.append([batch, "dc" if cur == "jacobi" else "jacobi"])
        table[str(n)] = ladder
    return table


def tune_eigh(lib, ns=EIGH_NS, batches=EIGH_BATCHES, reps=3):
    """Eigensolver ladders per dimension (eigh_ladders) and the raw timings."""
    from . import _native as N
    raw = {}
    for n in ns:
        for batch in batches:
            tj = time_eigh(lib, n, batch, N.MSP_EIGH_JACOBI, reps)
            td = time_eigh(lib, n, batch, N.MSP_EIGH_DC, reps)
            raw[f"{n}x{batch}"] = {"jacobi_ms": tj, "dc_ms": td}
    return eigh_ladders(raw), raw


def tune_timesum(lib, ns=TIMESUM_NS, ntimes=2000, batch=256, reps=3):
    """Smallest n (of ns) from which the mixed-precision time sum is faster
    than the FP64 one, and the raw timings."""
    import torch
    from . import _native as N
    raw = {}
    first = None
    st = torch.cuda.current_stream().cuda_stream
    t = torch.linspace(0, 10, ntimes, dtype=torch.float64, device="cuda")
    for n in ns:
        E = torch.sort(torch.randn((batch, n), dtype=torch.float64, device="cuda") * 50, dim=1).values
        C = torch.randn((batch, n, n), dtype=torch.float64, device="cuda")
        C = ((C + C.transpose(1, 2)) / 2).contiguous()
        out = torch.empty((batch, ntimes), dtype=torch.float64, device="cuda")
        ws = lib.msp_timesum_mixed_workspace(n, batch, ntimes)
        work = torch.empty((max(ws, 8) // 8,), dtype=torch.float64, device="cuda")
        t64 = _events_ms(lambda: N.check(lib.msp_timesum(n, batch, ntimes, E.data_ptr(), C.data_ptr(),
                                                         1, t.data_ptr(), None, 0, out.data_ptr(),
                                                         st)), reps)
        t32 = _events_ms(lambda: N.check(lib.msp_timesum_mixed(n, batch, ntimes, E.data_ptr(),
                                                               C.data_ptr(), t.data_ptr(), None,
                                                               out.data_ptr(), work.data_ptr(), ws,
                                                               st)), reps)
        raw[str(n)] = {"fp64_ms": t64, "mixed_ms": t32}
        if first is None and t32 < t64:
            first = n
    return (first if first is not None else 1 << 30), raw


def tune_timesum64(lib, ns=TIMESUM_NS + (192, 256), ntimes=2000, batch=256, reps=3, tol=0.05):
    """Smallest n (of ns) from which the FP64 matrix-core time sum (phase
    table + MFMA GEMM) is faster than the FP64 VALU one by more than `tol`
    at this n and every larger n measured, and the raw timings."""
    import torch
    from . import _native as N
    raw = {}
    st = torch.cuda.current_stream().cuda_stream
    t = torch.linspace(0, 10, ntimes, dtype=torch.float64, device="cuda")
    for n in ns:
        E = torch.sort(torch.randn((batch, n), dtype=torch.float64, device="cuda") * 50, dim=1).values
        C = torch.randn((batch, n, n), dtype=torch.float64, device="cuda")
        C = ((C + C.transpose(1, 2)) / 2).contiguous()
        out = torch.empty((batch, ntimes), dtype=torch.float64, device="cuda")
        ws = lib.msp_timesum_mfma64_workspace(n, batch, ntimes)
        work = torch.empty((max(ws, 8) // 8,), dtype=torch.float64, device="cuda")
        t64 = _events_ms(lambda: N.check(lib.msp_timesum(n, batch, ntimes, E.data_ptr(), C.data_ptr(),
                                                         1, t.data_ptr(), None, 0, out.data_ptr(),
                                                         st)), reps)
        tm = _events_ms(lambda: N.check(lib.msp_timesum_mfma64(n, batch, ntimes, E.data_ptr(),
                                                               C.data_ptr(), t.data_ptr(), None,
                                                               out.data_ptr(), work.data_ptr(), ws,
                                                               st)), reps)
        raw[str(n)] = {"valu_ms": t64, "mfma_ms": tm}
        del E, C, out, work
        torch.cuda.empty_cache()
    return timesum64_threshold(raw, tol), raw


def timesum64_threshold(raw, tol=0.05):
    """The smallest n from which every measured n has the matrix-core time
    sum faster by more than `tol` (a very large n when none does)."""
    ns = sorted(int(k) for k in raw)
    first = 1 << 30
    for n in reversed(ns):
        r = raw[str(n)]
        if r["mfma_ms"] < r["valu_ms"] * (1.0 - tol):
            first = n
        else:
            break
    return first


def tune_chunk(candidates=(2048, 4096, 8192), nfields=40, nori=500):
    """Instances per launch for the n = 192 integral sweep: best of one
    timed execute per candidate on a 40-field slice of the sweep."""
    import torch
    from . import workloads as W
    from .engine import SpinEngine
    systems, batch = W.c2_alc_sweep(nfields=nfields, nori=nori)
    raw = {}
    for c in candidates:
        eng = SpinEngine("cuda:0", max_chunk=c, tuning=None)
        plan = eng.prepare(systems, batch)
        eng.execute(plan)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.execute(plan)
        torch.cuda.synchronize()
        raw[str(c)] = (time.perf_counter() - t0) * 1e3
        del plan
        torch.cuda.empty_cache()
    return choose_chunk(raw), raw


def choose_chunk(raw, tol=0.02):
    """The largest candidate within `tol` of the fastest: larger chunks fill
    whole residency waves of the full sweep, which the slice does not show."""
    fastest = min(raw.values())
    return max(int(c) for c, ms in raw.items() if ms <= fastest * (1 + tol))


def tune(quick=False):
    import torch
    from . import _native as N
    lib = N.lib()
    reps = 2 if quick else 3
    batches = (64, 1024, 8192) if quick else EIGH_BATCHES
    eigh, eigh_raw = tune_eigh(lib, batches=batches, reps=reps)
    ts_min, ts_raw = tune_timesum(lib, reps=reps)
    ts64_min, ts64_raw = tune_timesum64(lib, reps=reps)
    chunk, chunk_raw = tune_chunk() if not quick else (8192, {})
    return {
        "device": torch.cuda.get_device_properties(0).gcnArchName,
        "eigh": eigh,
        "mixed_timesum_min_n": ts_min,
        "fp64_timesum_mfma_min_n": ts64_min,
        "max_chunk": chunk,
        "measured": {"eigh_ms": eigh_raw, "timesum_ms": ts_raw, "timesum64_ms": ts64_raw,
                     "chunk_ms": chunk_raw},
    }


def load(path=DEFAULT_PATH):
    """The tuning dict at `path`, or None when there is none."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def eigh_method(tuning, n, batch):
    """MSP_EIGH_* for n instances of dimension n from a tuning dict (AUTO
    when the tuning has no entry for n)."""
    from . import _native as N
    if not tuning:
        return N.MSP_EIGH_AUTO
    ladder = tuning.get("eigh", {}).get(str(int(n)))
    if not ladder:
        return N.MSP_EIGH_AUTO
    choice = ladder[0][1]
    for bmin, m in ladder:
        if batch >= bmin:
            choice = m
    return N.MSP_EIGH_JACOBI if choice == "jacobi" else N.MSP_EIGH_DC


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--quick", action="store_true", help="fewer batch sizes, no chunk sweep")
    args = ap.parse_args(argv)
    res = tune(quick=args.quick)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print(json.dumps({k: v for k, v in res.items() if k != "measured"}), flush=True)


if __name__ == "__main__":
    main()
