"""Adjoint gradients of the powder-averaged fast-path polarisation with
respect to the parameters of the field-independent Hamiltonian, for fitting.

The reference fits by finite differences (muspinsim/fitting.py:99-105), one
full simulation per parameter and per gradient; here one forward pass and
one adjoint pass give the whole gradient.  The adjoint pass recomputes the
eigenbasis of every chunk instead of storing it (H build, eigh, O = U^H
sigma U), forms the eigenbasis gradient G' with msp_adjoint_freq /
msp_hadamard_real / msp_gemm_cn / msp_adjoint_gprime, rotates it back,
G = U G' U^H, sums it over instances and contracts it with each dH/dtheta_p
(msp_trace_products).  Degenerate levels (zero field) use the analytic limit
of the divided differences, so the gradient is defined there.

`PolarisationModel` wraps a parametrised system; `polarisation` is a
torch.autograd.Function entry point: P = polarisation(theta, model) is the
results array, and backward() returns dL/dtheta for any upstream dL/dP.
"""

import numpy as np
import torch

from . import _native as N
from . import gpu
from .engine import SystemDesc


class PolarisationModel:
    """build(theta) -> SystemDesc; dterms(theta) -> one term list per parameter
    with the derivative tensors dT/dtheta_p (terms are linear in their
    tensors, so dH_p is the Hamiltonian of those terms); batch: an
    InstanceBatch that every instance of which takes the fast route
    (T = inf or B = 0, muon first, asymmetry)."""

    def __init__(self, build, dterms, batch, engine=None, degtol=1e-8):
        self.build = build
        self.dterms = dterms
        self.batch = batch
        self.engine = engine or gpu.engine()
        self.degtol = degtol

    def _plan(self, theta):
        system = self.build(np.asarray(theta, dtype=float))
        plan = self.engine.prepare([system], self.batch)
        for g in plan.groups:
            if g.route != "fast":
                raise NotImplementedError("the adjoint covers the fast route (T = inf or B = 0, "
                                          "muon first, asymmetry)")
        return system, plan

    def forward(self, theta):
        """Results array (n_slots, T) on the device."""
        _, plan = self._plan(theta)
        res = self.engine.execute(plan).clone()
        self.engine.check_converged(plan)
        return res

    def dH(self, theta, system):
        """(nparam, n, n) complex derivative Hamiltonians on the device."""
        mats = []
        for terms in self.dterms(np.asarray(theta, dtype=float)):
            desc = SystemDesc(Is=list(system.Is), gammas=list(system.gammas), terms=list(terms),
                              muon_index=system.muon_index)
            mats.append(gpu.build_h(desc))
        return gpu.to_dev(np.array(mats))

    def vjp(self, theta, grad):
        """dL/dtheta (numpy, nparam) for upstream grad = dL/dresults (n_slots, T)."""
        eng = self.engine
        L = eng.lib
        st = eng.stream()
        system, plan = self._plan(theta)
        batch = plan.batch
        grad = torch.as_tensor(grad, dtype=torch.float64, device=eng.device).contiguous()
        n = plan.groups[0].n if plan.groups else system.n
        Gsum = torch.zeros((n, n), dtype=torch.complex128, device=eng.device)
        slots = torch.as_tensor(batch.slot.astype(np.int32), device=eng.device)
        for g in plan.groups:
            if g.base is not None:
                N.check(L.msp_build_h(n, g.base.shape[0], g.base_sys.data_ptr(),
                                      plan.d_terms.data_ptr(), g.base_ids.data_ptr(), None, None, 0,
                                      g.base.data_ptr(), st))
            gslot = slots[g.d_idx]
            for c0 in range(0, g.idx.size, g.cs):
                nb = min(g.cs, g.idx.size - c0)
                sl = slice(c0, c0 + nb)

                def p(t):
                    return t[sl].data_ptr()

                N.check(L.msp_rotate_fields(nb, p(g.q), p(g.B), p(g.Bint), p(g.p),
                                            g.Br.data_ptr(), g.pr.data_ptr(), st))
                if g.base is not None:
                    N.check(L.msp_build_h(n, nb, g.base_sys.data_ptr(), plan.d_terms.data_ptr(),
                                          p(g.sid_local), g.Br.data_ptr(), g.base.data_ptr(), 1,
                                          g.H.data_ptr(), st))
                else:
                    N.check(L.msp_build_h(n, nb, plan.d_sys.data_ptr(), plan.d_terms.data_ptr(),
                                          p(g.sid), g.Br.data_ptr(), None, 1, g.H.data_ptr(), st))
                N.check(L.msp_eigh(n, nb, g.H.data_ptr(), g.E.data_ptr(), g.U.data_ptr(),
                                   g.info.data_ptr(), eng.method, g.work.data_ptr(), g.wsb, st))
                Y = g.H
                N.check(L.msp_apply_muon_op(n, nb, plan.d_sys.data_ptr(), p(g.sid),
                                            g.pr.data_ptr(), 1, g.U.data_ptr(), Y.data_ptr(), st))
                O = g.O[:nb]
                N.check(L.msp_gemm_cn(n, nb, g.U.data_ptr(), Y.data_ptr(), O.data_ptr(), st))
                R = torch.empty((nb, n, n), dtype=torch.float64, device=eng.device)
                S = torch.empty_like(R)
                sl_slot = gslot[sl].contiguous()
                N.check(L.msp_adjoint_freq(n, nb, plan.width, g.E.data_ptr(),
                                           plan.times.data_ptr(), grad.data_ptr(),
                                           sl_slot.data_ptr(), p(g.w), 1.0 / batch.avg_N,
                                           R.data_ptr(), S.data_ptr(), st))
                M = torch.empty((nb, n, n), dtype=torch.complex128, device=eng.device)
                N.check(L.msp_hadamard_real(n, nb, R.data_ptr(), O.data_ptr(), M.data_ptr(), st))
                MO = eng.gemm_cn(M, O)
                OM = eng.gemm_cn(O, M)
                Gp = torch.empty_like(M)
                N.check(L.msp_adjoint_gprime(n, nb, g.E.data_ptr(), O.data_ptr(), MO.data_ptr(),
                                             OM.data_ptr(), S.data_ptr(), float(g.d_other),
                                             float(self.degtol), Gp.data_ptr(), st))
                U = g.U[:nb]
                Uh = U.conj().transpose(-1, -2).contiguous()
                X = eng.gemm_cn(Uh, Gp)                          # U G'
                Xh = X.conj().transpose(-1, -2).contiguous()
                G = eng.gemm_cn(Xh, Uh)                          # U G' U^H
                N.check(L.msp_batch_sum(n * n, nb, G.data_ptr(), Gsum.data_ptr(), st))
                plan.checks.append(g.info[:nb].amax())
        self.engine.check_converged(plan)
        dH = self.dH(theta, system)
        out = torch.empty((dH.shape[0],), dtype=torch.float64, device=eng.device)
        N.check(L.msp_trace_products(n, dH.shape[0], Gsum.data_ptr(), dH.data_ptr(),
                                     out.data_ptr(), st))
        return out.cpu().numpy()


class _Polarisation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model):
        ctx.model = model
        ctx.theta = theta.detach().cpu().numpy().astype(float)
        return model.forward(ctx.theta)

    @staticmethod
    def backward(ctx, grad):
        g = ctx.model.vjp(ctx.theta, grad)
        return torch.as_tensor(g, dtype=torch.float64), None


def polarisation(theta, model):
    """Differentiable results array (n_slots, T) of model at parameters theta
    (a float64 CPU tensor with requires_grad)."""
    return _Polarisation.apply(theta, model)


def fit(model, target, theta0, sigma=1.0, bounds=None, tol=1e-10, maxiter=200):
    """Least-squares fit of the model to target (n_slots, T) by L-BFGS-B on
    chi^2 = sum ((P - target) / sigma)^2 with the adjoint gradient: one forward
    and one adjoint pass per iteration instead of the reference's
    (n_params + 1) simulations per finite-difference gradient
    (muspinsim/fitting.py:99-105)."""
    from scipy.optimize import minimize
    target = np.asarray(target, dtype=float)

    def fg(theta):
        P = model.forward(theta).cpu().numpy()
        r = (P - target) / sigma ** 2
        chi2 = float(np.sum((P - target) ** 2) / sigma ** 2)
        return chi2, model.vjp(theta, 2.0 * r)

    return minimize(fg, np.asarray(theta0, dtype=float), jac=True, method="L-BFGS-B",
                    bounds=bounds, tol=tol, options={"maxiter": maxiter})
