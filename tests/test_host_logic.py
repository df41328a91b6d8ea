"""Host-side preparation of the batched path (no GPU): per-instance inputs of
the dissipative route (vec(rho0), vec(O^T), jump operators, rates) and of
the finite-temperature route against the oracle's restatement of
ExperimentRunner.rho0 / dissipation_operators (experiment.py:169-325)."""

import numpy as np
import pytest

from oracle import physics
from oracle import pipeline as opipe


def _batch(rng, nb, B, T):
    from muspinsim_amd.engine import InstanceBatch
    q = rng.normal(size=(nb, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    p = rng.normal(size=(nb, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    return InstanceBatch(q=q, B=np.tile(np.asarray(B, dtype=float), (nb, 1)),
                         Bint=rng.normal(size=(nb, 3)) * 0.001 if np.any(B) else np.zeros((nb, 3)),
                         p=p, T=np.full(nb, T), w=np.ones(nb), slot=np.zeros(nb, np.int64),
                         sys_id=np.zeros(nb, np.int64), times=np.zeros(1), y="asymmetry",
                         n_slots=1, avg_N=nb)


@pytest.mark.parametrize("B,T", [((0, 0, 0), np.inf), ((0, 0, 0.1), 2.0), ((0.3, -0.2, 1.0), 0.5),
                                 ((0, 0, -1.0), 0.0)])
def test_lindblad_inputs_match_oracle(B, T):
    from muspinsim_amd.constants import ELEC_GAMMA, MU_GAMMA
    from muspinsim_amd.engine import SystemDesc, lindblad_inputs_host, rotate_host
    rng = np.random.default_rng(3)
    Is = [0.5, 0.5, 1.0]
    g = [MU_GAMMA, ELEC_GAMMA, 3.0777]
    systems = [SystemDesc(Is, g, [], 0, dissip={1: 7.0, 2: 0.5}),
               SystemDesc(Is, g, [], 0, dissip={1: 70.0, 2: 3.5})]
    batch = _batch(rng, 5, B, T)
    batch.sys_id = np.array([0, 1, 1, 0, 1])
    idx = np.arange(5)
    Br = rotate_host(batch.q, batch.B) + batch.Bint
    r, u, A, rates = lindblad_inputs_host(systems, batch, idx, Br)
    for k in idx:
        s = systems[batch.sys_id[k]]
        Brk, prk, _ = opipe.load_config(batch.q[k], batch.B[k], batch.Bint[k], batch.p[k], T, 0)
        assert np.abs(Br[k] - Brk).max() < 1e-14
        rho = physics.rho0(Is, g, 0, Brk, prk, T)
        assert np.abs(r[k] - rho.reshape(-1)).max() < 1e-14
        O = physics.muon_operator(Is, 0, prk)
        assert np.abs(u[k, 0] - O.T.reshape(-1)).max() < 1e-14
        dops = physics.dissipation_operators(Is, g, Brk, T, s.dissip)
        assert A.shape[1] == len(dops)
        for j, (Aj, gj) in enumerate(dops):
            assert np.abs(A[k, j] - Aj).max() < 1e-14
            assert abs(rates[k, j] - gj) <= 1e-14 * max(1.0, abs(gj))


def test_c4_dissipative_builder():
    from muspinsim_amd import workloads as W
    systems, batch = W.c4_dissipative(ninst=16, ntimes=50)
    assert len(systems) == 16 and len(batch) == 16
    assert all(s.n == 8 and list(s.dissip) == [1] for s in systems)
    assert all(1.0 <= s.dissip[1] <= 100.0 for s in systems)
    assert batch.times.shape == (50,) and batch.avg_N == 16


def test_autotune_ladder_and_engine_choice():
    """The tuned eigensolver ladder (batch thresholds per dimension) and the
    engine's per-group choice: explicit methods win, Jacobi is never chosen
    above n = 64, missing entries fall back to AUTO."""
    from muspinsim_amd import _native as N
    from muspinsim_amd import autotune
    t = {"eigh": {"64": [[1, "jacobi"], [1024, "dc"]], "16": [[1, "jacobi"]]}}
    assert autotune.eigh_method(t, 64, 256) == N.MSP_EIGH_JACOBI
    assert autotune.eigh_method(t, 64, 1024) == N.MSP_EIGH_DC
    assert autotune.eigh_method(t, 64, 50000) == N.MSP_EIGH_DC
    assert autotune.eigh_method(t, 16, 10 ** 6) == N.MSP_EIGH_JACOBI
    assert autotune.eigh_method(t, 32, 10) == N.MSP_EIGH_AUTO
    assert autotune.eigh_method(None, 64, 10) == N.MSP_EIGH_AUTO
    assert autotune.load("/nonexistent/tuning.json") is None


def test_autotune_ladder_hysteresis():
    """Ladders from raw timings: the faster method from each batch on, but a
    difference within the tolerance keeps the previous choice."""
    from muspinsim_amd import autotune
    raw = {"12x64": {"jacobi_ms": 0.116, "dc_ms": 0.118},
           "12x256": {"jacobi_ms": 0.121, "dc_ms": 0.119},     # 1.7 % apart: stays Jacobi
           "12x1024": {"jacobi_ms": 0.137, "dc_ms": 0.187},
           "16x64": {"jacobi_ms": 0.165, "dc_ms": 0.135},
           "16x1024": {"jacobi_ms": 0.188, "dc_ms": 0.230},    # Jacobi 18 % faster: switch
           "16x4096": {"jacobi_ms": 0.401, "dc_ms": 0.390}}    # 2.7 % apart: stays Jacobi
    t = autotune.eigh_ladders(raw, tol=0.05)
    assert t["12"] == [[1, "jacobi"]]
    assert t["16"] == [[1, "dc"], [1024, "jacobi"]]
    assert autotune.eigh_ladders(raw, tol=0.0)["12"] == [[1, "jacobi"], [256, "dc"], [1024, "jacobi"]]
