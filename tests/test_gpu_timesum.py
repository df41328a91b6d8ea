"""Time sums of the real routes on matrix cores, on the GPU.

msp_timesum_mfma64 (FP64 phase table + FP64 MFMA GEMM) forms the same phases
as msp_timesum with the same expression and differs from it only in the
order of the FP64 sums: both are held to 1e-13 sum_ij |C_ij| of each other
and to the numpy sum.  End to end, SpinEngine is forced onto each time-sum
path (thresholds 0) and compared with the FP64 oracle: <= 1e-10 (PARITY) for
the FP64 paths, <= 1e-6 for the mixed one.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()


def _timesum_inputs(rng, n, batch, T, spread=2000.0, tmax=20.0):
    E = np.sort(rng.normal(size=(batch, n)) * spread, axis=1)
    G = rng.normal(size=(batch, n, n))
    C = np.ascontiguousarray((G + np.swapaxes(G, 1, 2)) / (2 * n * n))
    t = np.linspace(0, tmax, T)
    w = rng.uniform(0.5, 2.0, size=batch)
    return E, C, t, w


@pytest.mark.parametrize("n,T,batch", [(1, 5, 2), (16, 64, 3), (33, 70, 11), (64, 200, 4),
                                       (100, 129, 5), (512, 300, 2)])
def test_timesum_mfma64_vs_fp64_valu(gpu_lib, n, T, batch):
    """FP64 matrix-core time sum against msp_timesum (FP64 VALU) and numpy,
    including row, column and time tails and phases up to ~1e6 rad."""
    from muspinsim_amd import _native as N
    rng = np.random.default_rng(100 * n + T)
    E, C, t, w = _timesum_inputs(rng, n, batch, T)
    dE, dC, dt, dw = (_dev(x) for x in (E, C, t, w))
    out = torch.full((batch, T), np.nan, dtype=torch.float64, device="cuda")
    ws = gpu_lib.msp_timesum_mfma64_workspace(n, batch, T)
    work = torch.empty((ws // 8,), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    N.check(gpu_lib.msp_timesum_mfma64(n, batch, T, dE.data_ptr(), dC.data_ptr(), dt.data_ptr(),
                                       dw.data_ptr(), out.data_ptr(), work.data_ptr(), ws, st))
    valu = torch.empty_like(out)
    N.check(gpu_lib.msp_timesum(n, batch, T, dE.data_ptr(), dC.data_ptr(), 1, dt.data_ptr(),
                                dw.data_ptr(), 0, valu.data_ptr(), st))
    got, g64 = out.cpu().numpy(), valu.cpu().numpy()
    assert np.all(np.isfinite(got))
    for b in range(batch):
        scale = w[b] * np.abs(C[b]).sum()
        assert np.abs(got[b] - g64[b]).max() <= 1e-13 * scale
        ph = np.cos(2 * np.pi * (E[b][:, None] - E[b][None, :])[None] * t[:, None, None])
        ref = w[b] * np.einsum("tij,ij->t", ph, C[b])
        phmax = 2 * np.pi * (E[b].max() - E[b].min()) * t.max()
        assert np.abs(got[b] - ref).max() <= 4e-16 * max(phmax, 1.0) * scale


def test_timesum_mfma64_argument_errors(gpu_lib):
    """Workspace too small, batch above the grid limit and null pointers are
    value errors with the entry point named; empty sizes are no-ops."""
    from muspinsim_amd import _native as N
    n, batch, T = 8, 2, 10
    E = _dev(np.zeros((batch, n)))
    C = _dev(np.zeros((batch, n, n)))
    t = _dev(np.zeros(T))
    out = torch.empty((batch, T), dtype=torch.float64, device="cuda")
    ws = gpu_lib.msp_timesum_mfma64_workspace(n, batch, T)
    assert ws == 2 * n * 64 * 8 * batch
    work = torch.empty((ws // 8,), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="msp_timesum_mfma64: workspace"):
        N.check(gpu_lib.msp_timesum_mfma64(n, batch, T, E.data_ptr(), C.data_ptr(), t.data_ptr(),
                                           None, out.data_ptr(), work.data_ptr(), ws - 8, None))
    with pytest.raises(ValueError, match="batch > 65535"):
        N.check(gpu_lib.msp_timesum_mfma64(n, 65536, T, E.data_ptr(), C.data_ptr(), t.data_ptr(),
                                           None, out.data_ptr(), work.data_ptr(), ws, None))
    with pytest.raises(ValueError, match="null pointer"):
        N.check(gpu_lib.msp_timesum_mfma64(n, batch, T, None, C.data_ptr(), t.data_ptr(), None,
                                           out.data_ptr(), work.data_ptr(), ws, None))
    N.check(gpu_lib.msp_timesum_mfma64(n, 0, T, None, None, None, None, None, None, 0, None))


@pytest.mark.parametrize("precision,cfg", [("fp64", "c3"), ("fp64", "c5"), ("fp64", "c1"),
                                           ("mixed", "c3"), ("mixed", "c5")])
def test_engine_timesum_paths_vs_oracle(gpu_lib, precision, cfg):
    """SpinEngine with both time-sum thresholds at 0, so that every real-route
    group of the reduced configuration runs the matrix-core time sum of its
    precision; against the FP64 oracle."""
    from muspinsim_amd import workloads as W
    from muspinsim_amd.engine import SpinEngine
    from oracle import pipeline as opipe
    systems, batch = {"c3": lambda: W.c3_tf_stress(nori=3, ntimes=400),
                      "c5": lambda: W.c5_fit(nori=16, ntimes=500),
                      "c1": lambda: W.c1_zero_field_powder(nori=20, ntimes=300)}[cfg]()
    tuning = {"mixed_timesum_min_n": 1, "mixed_timesum_min_work": 0.0,
              "fp64_timesum_mfma_min_n": 1}
    eng = SpinEngine("cuda:0", precision=precision, tuning=tuning)
    plan = eng.prepare(systems, batch)
    got = eng.execute(plan).cpu().numpy()
    eng.check_converged(plan)
    want = "mixed" if precision == "mixed" else "mfma64"
    real = [g for g in plan.groups if g.route in ("fast", "iso")]
    assert real and all(g.tsmode == want for g in real)
    ref = opipe.run(systems, batch)
    err = np.abs(got - ref).max()
    assert err <= (1e-6 if precision == "mixed" else 1e-10), f"max |dP| = {err:.3e}"
