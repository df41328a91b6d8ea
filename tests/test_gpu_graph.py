"""Graph capture and replay: the msp_graph_* C-ABI helpers around a raw
msp_eigh sequence, and SpinEngine.execute_graph on whole plans (divide and
conquer with its side stream, Jacobi, and the thermal route).  Every replay
runs after unrelated eager work and a synchronisation and must reproduce the
eager results bit for bit (same kernels, same inputs)."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _herm(rng, n):
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (X + X.conj().T) / 2


def test_capi_graph_replays_eigh(gpu_lib):
    from muspinsim_amd import _native as N
    L = N.lib()
    n, B = 192, 6
    rng = np.random.default_rng(3)
    H0 = torch.as_tensor(np.array([_herm(rng, n) for _ in range(B)])).cuda()
    H = torch.empty_like(H0)
    E = torch.empty((B, n), dtype=torch.float64, device="cuda")
    U = torch.empty_like(H0)
    info = torch.empty((B,), dtype=torch.int32, device="cuda")
    wsb = L.msp_eigh_workspace(n, B, N.MSP_EIGH_DC)
    work = torch.empty((wsb // 8 + 1,), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    sp = s.cuda_stream

    def seq():
        H.copy_(H0)
        N.check(L.msp_eigh(n, B, H.data_ptr(), E.data_ptr(), U.data_ptr(), info.data_ptr(),
                           N.MSP_EIGH_DC, work.data_ptr(), wsb, sp))

    with torch.cuda.stream(s):
        seq()                                    # eager run first
    torch.cuda.synchronize()
    E0, U0 = E.clone(), U.clone()
    gx = ctypes.c_void_p()
    with torch.cuda.stream(s):
        N.check(L.msp_graph_begin(sp))
        seq()
        N.check(L.msp_graph_end(sp, ctypes.byref(gx)))
    try:
        other = torch.as_tensor(np.array([_herm(rng, 64) for _ in range(4)])).cuda()
        for rep in range(3):
            E.fill_(float("nan"))
            U.zero_()
            from muspinsim_amd import gpu
            gpu.engine().eigh(other.clone())     # unrelated eager work
            torch.cuda.synchronize()
            N.check(L.msp_graph_launch(gx, sp))
            torch.cuda.synchronize()
            assert torch.equal(E, E0) and torch.equal(U, U0), f"replay {rep}"
            assert int(info.max()) == 0
    finally:
        N.check(L.msp_graph_destroy(gx))


@pytest.mark.parametrize("case", ["alc_n192", "zero_field_n16", "thermal"])
def test_engine_graph_replays_plan(gpu_lib, case):
    from muspinsim_amd import workloads as W
    from muspinsim_amd.engine import SpinEngine
    if case == "alc_n192":
        systems, batch = W.c2_alc_sweep(nfields=6, nori=3)
    elif case == "zero_field_n16":
        systems, batch = W.c1_zero_field_powder(nori=20, ntimes=100)
    else:
        systems, batch = W.c4_uneven_batch(seed=7, ninst=24, ntimes=21, max_dim=96)
        batch.T = np.full(batch.w.size, 15.0)
    eng = SpinEngine("cuda:0", max_chunk=7)      # several chunks per group
    plan = eng.prepare(systems, batch)
    ref = eng.execute(plan).clone()
    torch.cuda.synchronize()
    other_sys, other_batch = W.c1_zero_field_powder(nori=5, ntimes=50)
    for rep in range(3):
        eng.run(other_sys, other_batch)          # unrelated eager work
        torch.cuda.synchronize()
        res = eng.execute_graph(plan)
        torch.cuda.synchronize()
        assert torch.equal(res, ref), f"replay {rep}"
        eng.check_converged(plan)
