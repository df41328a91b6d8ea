"""Adjoint gradient of the powder-averaged polarisation (fitting): the
parameter derivatives of the c5 system against finite differences of H
(CPU), and the GPU vector-Jacobian product against central finite
differences of the GPU forward and of the FP64 oracle forward, including a
zero-field case with degenerate levels.

Tolerances: dH 1e-6 relative (central differences, h = 1e-4 relative);
gradients 1e-5 relative to the gradient norm, against the fourth-order
central difference (f(-2h) - 8 f(-h) + 8 f(h) - f(2h)) / 12h with
h = 1e-6 |theta_k|: a hyperfine step of 1e-4 relative would turn the phase
of a 300 MHz line by ~2 rad over 10 us, so larger steps measure curvature.
"""

import numpy as np
import pytest

from oracle import physics
from oracle import pipeline as opipe


def _h_of(params, seed=5):
    from muspinsim_amd import workloads as W
    s, _ = W.c5_system(params, seed)
    return physics.hamiltonian(s.Is, s.terms)


def test_c5_dterms_match_finite_differences():
    from muspinsim_amd import workloads as W
    s, p0 = W.c5_system(None)
    dts = W.c5_dterms(p0)
    assert len(dts) == len(W.C5_PARAMS) == 10
    for k, terms in enumerate(dts):
        h = 1e-4 * abs(p0[k])
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        fd = (_h_of(pp) - _h_of(pm)) / (2 * h)
        an = physics.hamiltonian(s.Is, terms)
        assert np.abs(fd - an).max() <= 1e-6 * max(1.0, np.abs(an).max())


def _model(nori, ntimes, B=(0.0, 0.0, 0.0)):
    from muspinsim_amd import workloads as W
    from muspinsim_amd.adjoint import PolarisationModel
    times = np.linspace(0.0, 10.0, ntimes)
    batch = W.powder_batch(nori, [list(B)], times, "asymmetry", [0, 0, 1.0])
    return PolarisationModel(lambda th: W.c5_system(th)[0], W.c5_dterms, batch), batch


def _fd4(f, theta, k, rel=1e-6):
    h = rel * abs(theta[k])
    vals = []
    for m in (-2, -1, 1, 2):
        th = theta.copy()
        th[k] += m * h
        vals.append(f(th))
    return (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)


def _chi2_and_grad(model, theta, target, sigma):
    P = model.forward(theta).cpu().numpy()
    r = 2.0 * (P - target) / sigma ** 2
    return float(np.sum((P - target) ** 2) / sigma ** 2), model.vjp(theta, r)


@pytest.mark.gpu
def test_adjoint_matches_gpu_finite_differences(gpu_lib):
    from muspinsim_amd import workloads as W
    model, _ = _model(nori=8, ntimes=300)
    _, p0 = W.c5_system(None)
    rng = np.random.default_rng(0)
    target = model.forward(p0 * (1 + 0.01 * rng.normal(size=p0.size))).cpu().numpy()
    target = target + rng.normal(0, 0.005, size=target.shape)
    chi2, g = _chi2_and_grad(model, p0, target, 0.005)

    def f(th):
        P = model.forward(th).cpu().numpy()
        return float(np.sum((P - target) ** 2) / 0.005 ** 2)

    fd = np.array([_fd4(f, p0, k) for k in range(p0.size)])
    assert np.abs(g - fd).max() <= 1e-5 * np.abs(fd).max()


@pytest.mark.gpu
def test_adjoint_matches_oracle_finite_differences(gpu_lib):
    """Reference gradient from the FP64 oracle forward (central differences)
    on a small zero-field powder, and the GPU vjp of the same linear
    functional."""
    from muspinsim_amd import workloads as W
    model, batch = _model(nori=4, ntimes=120)
    _, p0 = W.c5_system(None)
    r = np.cos(np.linspace(0, 3, 120))[None, :]

    def oracle_L(th):
        s, _ = W.c5_system(th)
        return float(np.sum(opipe.run([s], batch) * r))

    g = model.vjp(p0, r)
    fd = np.array([_fd4(oracle_L, p0, k) for k in range(p0.size)])
    assert np.abs(g - fd).max() <= 1e-5 * np.abs(fd).max()


@pytest.mark.gpu
def test_adjoint_degenerate_zero_field(gpu_lib):
    """Isotropic hyperfine at zero field (multiplets of degenerate levels):
    the divided-difference limits keep the gradient equal to the finite
    difference one."""
    from muspinsim_amd.adjoint import PolarisationModel
    from muspinsim_amd.constants import ELEC_GAMMA, MU_GAMMA
    from muspinsim_amd.engine import SystemDesc
    from muspinsim_amd import workloads as W
    Is = [0.5, 0.5, 0.5]
    g = [MU_GAMMA, ELEC_GAMMA, W.G_H]

    def build(th):
        return SystemDesc(Is, g, [("bilinear", 0, 1, th[0] * np.eye(3)),
                                  ("bilinear", 2, 1, th[1] * np.eye(3))], 0)

    def dterms(th):
        return [[("bilinear", 0, 1, np.eye(3))], [("bilinear", 2, 1, np.eye(3))]]

    times = np.linspace(0, 5, 200)
    batch = W.powder_batch(3, [[0, 0, 0]], times, "asymmetry", [0, 0, 1.0])
    model = PolarisationModel(build, dterms, batch)
    th = np.array([120.0, 15.0])
    r = np.sin(np.linspace(0, 7, 200))[None, :]
    gr = model.vjp(th, r)

    def f(t2):
        return float(np.sum(model.forward(t2).cpu().numpy() * r))

    fd = np.array([_fd4(f, th, k) for k in range(2)])
    assert np.abs(gr - fd).max() <= 1e-5 * np.abs(fd).max()


@pytest.mark.gpu
def test_autograd_function(gpu_lib):
    import torch
    from muspinsim_amd import workloads as W
    from muspinsim_amd.adjoint import polarisation
    model, _ = _model(nori=4, ntimes=100)
    _, p0 = W.c5_system(None)
    theta = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    P = polarisation(theta, model)
    loss = (P ** 2).sum()
    loss.backward()
    ref = model.vjp(p0, 2.0 * P.detach().cpu().numpy())
    assert np.abs(theta.grad.numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.gpu
def test_adjoint_fit_recovers_parameters(gpu_lib):
    """L-BFGS-B with the adjoint gradient recovers perturbed hyperfine
    components and a distance from a noiseless synthetic target."""
    from muspinsim_amd import workloads as W
    from muspinsim_amd.adjoint import PolarisationModel, fit
    model, _ = _model(nori=6, ntimes=200)
    _, p_true = W.c5_system(None)
    target = model.forward(p_true).cpu().numpy()
    free = [0, 3, 6]
    start = p_true.copy()
    # inside the basin of the oscillatory chi^2: a 300 MHz line drifts by
    # ~0.4 rad over 10 us at 2e-5 relative
    start[free] *= np.array([1.00002, 1.002, 1.002])

    def build(th):
        p = p_true.copy()
        p[free] = th
        return W.c5_system(p)[0]

    def dterms(th):
        p = p_true.copy()
        p[free] = th
        d = W.c5_dterms(p)
        return [d[k] for k in free]

    sub = PolarisationModel(build, dterms, model.batch)
    sol = fit(sub, target, start[free], sigma=0.005,
              bounds=[(250.0, 350.0), (0.0, 20.0), (1.0, 4.0)])
    assert np.all(np.abs(sol.x - p_true[free]) <= 1e-5 * np.abs(p_true[free]))
