"""Multi-process CPU timing of the oracle (the CPU baseline of bench.py).

TEST/BENCH INFRASTRUCTURE.  Mirrors the reference's MPI decomposition
(muspinsim/experiment.py:369: worker r takes instances[r::P]; mpi.py:104-112:
partial results summed) with forked worker processes, each limited to one
BLAS thread.  Must run before the parent process initialises the GPU.
"""

import multiprocessing as mp
import os
import time

import numpy as np

_STATE = {}


def _work(args):
    r, P = args
    from threadpoolctl import threadpool_limits

    from . import pipeline
    systems, batch, idx = _STATE["systems"], _STATE["batch"], _STATE["idx"]
    mine = idx[r::P]
    with threadpool_limits(1):
        t0 = time.perf_counter()
        rows = pipeline.instance_rows(systems, batch, mine)
        dt = time.perf_counter() - t0
    return r, rows, dt


def run(systems, batch, idx, procs):
    """Oracle rows for instances idx using `procs` forked workers.
    Returns (rows in idx order, wall seconds, per-worker busy seconds)."""
    procs = max(1, min(procs, len(idx)))     # no idle workers on small samples
    _STATE.update(systems=systems, batch=batch, idx=np.asarray(idx))
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    if procs == 1:
        res = [_work((0, 1))]
    else:
        with ctx.Pool(procs) as pool:
            res = pool.map(_work, [(r, procs) for r in range(procs)])
    wall = time.perf_counter() - t0
    width = res[0][1].shape[1] if len(res[0][1]) else 1
    rows = np.zeros((len(idx), width))
    for r, rr, _ in res:
        rows[r::procs] = rr
    return rows, wall, [dt for _, _, dt in res]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """CPUs this process may run on (its affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_share():
    """Worker processes the baseline may use: the affinity set, limited to
    the job's CPU share where the host states one through OMP_NUM_THREADS
    (a node that hosts one job per GPU gives each job such a share, and the
    other cores belong to the other GPUs' jobs)."""
    n = host_cores()
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def host_info():
    """CPU model and counts recorded beside the baseline."""
    return {"model": cpu_model(), "logical_cpus": os.cpu_count(),
            "affinity_cpus": host_cores(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
