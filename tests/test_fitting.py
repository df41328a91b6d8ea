"""FittingRunner drop-in: variable handling (no GPU) and the reference's
exponential-decay recovery (muspinsim/tests/test_fitting.py:26-140) on the
GPU with the three optimisation methods."""

import numpy as np
import pytest


def _cfg(x, g):
    from muspinsim_amd.config import MuSpinConfig
    from muspinsim_amd.spinsys import MuonSpinSystem
    return MuSpinConfig(MuonSpinSystem(["mu"]), time=x, dissipation={0: g})


def test_fit_variables():
    from muspinsim_amd.fitting import FittingRunner
    data = np.array([[0, 0.5], [1, 0.5], [2, 0.5]])
    f = FittingRunner(_cfg, {"B": (1.0, 0.0, 2.0), "A": 0.0}, data)
    assert f._xnames == ("A", "B")
    assert np.all(f._x == [0.0, 1.0])
    assert np.all(f._xbounds[0] == [-np.inf, np.inf])
    assert np.all(f._xbounds[1] == [0.0, 2.0])
    with pytest.raises(ValueError):
        FittingRunner(_cfg, {}, data)
    with pytest.raises(ValueError):
        FittingRunner(_cfg, {"g": 1.0}, data, method="newton")


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["nelder-mead", "lbfgs", "least-squares"])
def test_fit_exponential_decay(gpu_lib, method):
    from muspinsim_amd.fitting import FittingRunner
    t = np.linspace(0, 10.0, 100)
    g = 0.2
    data = np.stack([t, 0.5 * np.exp(-g * t)], axis=1)
    f = FittingRunner(_cfg, {"g": 0.5}, data, method=method)
    sol = f.run()
    assert abs(sol.x[0] - g) < 5e-4
    assert f.n_evaluations > 1
