"""Levenberg-Marquardt with the reference's API and control flow
(fit/base.py:14-156, fit/lm.py:15-539).

The fitter is the reference's, step for step (damping scheme, geodesic probe
at d = 0.1, curvature gate, sub-step loop, convergence tests, best-state
result); the arithmetic runs on the GPU:

* forward model and Jacobian: the gfx950 render/convolve kernels, float32;
* J^T W J and J^T W r: matrix-core contraction with float64 sums;
* rpp = J^T((2/d)(W (m1 - m0)/d - W J h)): float64 sums;
* the damped solve h = A(L)^-1 g: float64 Cholesky on the device;
* chi^2 / ndf: float64 reduction.

The host reads back one scalar per sub-step (chi^2, as the reference's
``.item()`` does) and the Np-vector h.
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from . import AP_config
from . import _lib as L
from .errors import OptimizeStop

__all__ = ["BaseOptimizer", "LM"]


class BaseOptimizer:
    """fit/base.py:14-156."""

    def __init__(self, model, initial_state: Sequence = None, relative_tolerance: float = 1e-3,
                 fit_window=None, **kwargs):
        self.model = model
        self.verbose = kwargs.get("verbose", 0)
        self.fit_window = self.model.window if fit_window is None else (fit_window & self.model.window)
        if initial_state is None:
            self.model.initialize()
            initial_state = self.model.parameters.vector_representation()
        self.current_state = torch.as_tensor(initial_state, dtype=AP_config.ap_dtype).cpu().clone()
        self.max_iter = kwargs.get("max_iter", 100 * len(self.current_state))
        self.iteration = 0
        self.relative_tolerance = relative_tolerance
        self.lambda_history = []
        self.loss_history = []
        self.message = ""

    def chi2min(self):
        return np.nanmin(self.loss_history)

    def res(self):
        """fit/base.py:104-117: the state with the lowest chi^2."""
        N = np.isfinite(self.loss_history)
        if np.sum(N) == 0:
            AP_config.ap_logger.warning("Getting optimizer res with no real loss history, using current state")
            return self.current_state.detach().cpu().numpy()
        return np.array(self.lambda_history)[N][np.argmin(np.array(self.loss_history)[N])]

    def res_loss(self):
        N = np.isfinite(self.loss_history)
        return np.min(np.array(self.loss_history)[N])


class LM(BaseOptimizer):
    """fit/lm.py:15-539."""

    def __init__(self, model, initial_state: Sequence = None, max_iter: int = 100,
                 relative_tolerance: float = 1e-5, ndf=None, **kwargs):
        super().__init__(model, initial_state, max_iter=max_iter, relative_tolerance=relative_tolerance,
                         **kwargs)
        self.max_iter = max_iter
        self.max_step_iter = kwargs.get("max_step_iter", 10)
        self.curvature_limit = kwargs.get("curvature_limit", 1.0)
        self._Lup = kwargs.get("Lup", 11.0)
        self._Ldn = kwargs.get("Ldn", 9.0)
        self.L = kwargs.get("L0", 1.0)
        self.acceleration = kwargs.get("acceleration", 0.0)
        self.lib = L.load()
        tgt = self.model.target[self.fit_window]
        self.Y = tgt.flatten("data").to(torch.float32).contiguous()
        kW = kwargs.get("W", None)
        if kW is not None:
            self.W = torch.as_tensor(kW, dtype=torch.float32, device=self.Y.device).flatten()
        elif self.model.target.has_variance:
            self.W = tgt.flatten("weight").contiguous()
        else:
            self.W = None
        # mask: True/1 = ignore (lm.py:204-222)
        fit_mask = self.model.fit_mask().flatten()
        if torch.sum(fit_mask).item() == 0:
            fit_mask = None
        mask = None
        if self.model.target.has_mask:
            mask = tgt.flatten("mask")
            if fit_mask is not None:
                mask = mask | fit_mask
        elif fit_mask is not None:
            mask = fit_mask
        self.mask = None if mask is None else mask.to(torch.uint8).contiguous()
        npix = self.Y.numel()
        nkeep = npix if self.mask is None else npix - int(self.mask.sum().item())
        if self.mask is not None and nkeep == 0:
            raise OptimizeStop("No data to fit. All pixels are masked")
        self.ndf = max(1.0, nkeep - len(self.current_state)) if ndf is None else ndf
        self._covariance_matrix = None
        dev = self.Y.device
        self.np = len(self.current_state)
        # groups of Sersic galaxies on their own windows (+ a flat sky) beyond
        # the dense kernels' 16 parameters go block-sparse (group_fit.py);
        # sparse=True / False forces the choice
        sparse = kwargs.get("sparse", None)
        self._sparse = None
        if sparse or (sparse is None and self.np > 16):
            from .group_fit import SparseGroup
            reason = SparseGroup.unsupported(self.model, tgt.window)
            if reason is None:
                self._sparse = SparseGroup(self.model, tgt.window)
            elif sparse:
                raise OptimizeStop(f"block-sparse fit not possible: {reason}")
        if self._sparse is None and self.np > 16:
            raise OptimizeStop(f"{self.np} parameters: the dense LM takes at most 16; the block-sparse one "
                               f"takes groups of Sersic galaxies on windows plus a flat sky")
        self._ws_bytes = self.lib.apx_reduce_workspace_bytes(npix, max(self.np, 1))
        self._ws = torch.empty(self._ws_bytes, dtype=torch.uint8, device=dev)
        self._dbuf = torch.empty(1 + 2 * 64 + 64 * 64 + 64, dtype=torch.float64, device=dev)
        self._status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.n_forward = 0
        self.n_jacobian = 0

    # -- device building blocks ---------------------------------------------
    def forward(self, parameters):
        self.n_forward += 1
        if self._sparse is not None:
            return self._sparse.forward(parameters)
        img = self.model(parameters=parameters.clone(), as_representation=True)
        return img.flatten("data").contiguous()

    def jacobian(self, parameters):
        self.n_jacobian += 1
        if self._sparse is not None:
            return self._sparse.jacobian(parameters)
        return self.model.jacobian(parameters=parameters.clone(), as_representation=True).flatten("data").contiguous()

    def _chi2(self, Ypred) -> torch.Tensor:
        """lm.py:373-378 (float64 on the device)."""
        out = self._dbuf[:1]
        L.check(self.lib.apx_chi2(L.ptr(Ypred), L.ptr(self.Y), L.ptr(self.W), L.ptr(self.mask),
                                  self.Y.numel(), L.ptr(out), L.ptr(self._ws), self._ws_bytes, L.stream()))
        return out[0] / self.ndf

    def _normal_equations(self, J, Y0):
        if self._sparse is not None:
            return self._sparse.normal_equations(J, Y0, self.Y, self.W, self.mask)
        n = self.np
        hess = torch.empty((n, n), dtype=torch.float64, device=self.Y.device)
        grad = torch.empty(n, dtype=torch.float64, device=self.Y.device)
        L.check(self.lib.apx_normal_equations(L.ptr(J), L.ptr(Y0), L.ptr(self.Y), L.ptr(self.W),
                                              L.ptr(self.mask), self.Y.numel(), n, L.ptr(hess), L.ptr(grad),
                                              L.ptr(self._ws), self._ws_bytes, L.stream()))
        return hess, grad

    def _h(self, Lval, grad, hess) -> torch.Tensor:
        """lm.py:359-371: float64 damped solve on the device."""
        if self._sparse is not None:
            return self._sparse.solve(Lval, grad, hess)
        h = torch.empty_like(grad)
        L.check(self.lib.apx_lm_solve(L.ptr(hess), L.ptr(grad), self.np, float(Lval), L.ptr(h),
                                      L.ptr(self._status), L.stream()))
        return h

    def _rpp(self, J, Y0, Y1, h, d):
        if self._sparse is not None:
            return self._sparse.rpp(J, Y0, Y1, self.W, self.mask, h, d)
        out = torch.empty(self.np, dtype=torch.float64, device=self.Y.device)
        L.check(self.lib.apx_rpp(L.ptr(J), L.ptr(Y0), L.ptr(Y1), L.ptr(self.W), L.ptr(self.mask),
                                 self.Y.numel(), self.np, L.ptr(h), float(d), L.ptr(out), L.ptr(self._ws),
                                 self._ws_bytes, L.stream()))
        return out

    def Lup(self):
        self.L = min(1e9, self.L * self._Lup)

    def Ldn(self):
        self.L = max(1e-9, self.L / self._Ldn)

    # -- the algorithm (lm.py:248-357, 427-493) ------------------------------
    @torch.no_grad()
    def step(self, chi2):
        Y0 = self.forward(self.current_state)
        J = self.jacobian(self.current_state)
        self.hess, self.grad = self._normal_equations(J, Y0)
        init_chi2 = chi2
        nostep = True
        best = (torch.zeros_like(self.current_state), init_chi2, self.L)
        scarry_best = (None, init_chi2, self.L)
        direction = "none"
        d = 0.1
        for iteration in range(self.max_step_iter):
            if iteration > self.max_step_iter / 2 and self.L < 1e-3:
                self.L = 1.0
            h_dev = self._h(self.L, self.grad, self.hess)
            h = h_dev.cpu()
            Y1 = self.forward(self.current_state + d * h)
            rpp = self._rpp(J, Y0, Y1, h_dev, d)
            if self.L > 1e-4:
                a = (-self._h(self.L, rpp, self.hess) / 2).cpu()
            else:
                a = torch.zeros_like(h)
            ha = h + a * self.acceleration
            Y1 = self.forward(self.current_state + ha)
            chi2 = self._chi2(Y1).item()
            if self.verbose > 1:
                AP_config.ap_logger.info(f"sub step L: {self.L}, Chi^2/DoF: {chi2}")
            if not np.isfinite(chi2):
                self.Lup()
                if direction == "better":
                    break
                direction = "worse"
                continue
            if chi2 <= scarry_best[1]:
                scarry_best = (ha, chi2, self.L)
            rho = torch.linalg.norm(a) / torch.linalg.norm(h)
            if rho > self.curvature_limit:
                self.Lup()
                if direction == "better":
                    break
                direction = "worse"
                continue
            if chi2 < best[1]:
                best = (ha, chi2, self.L)
                nostep = False
                self.Ldn()
                if self.L <= 1e-8 or direction == "worse":
                    break
                direction = "better"
            elif chi2 > best[1] and direction in ["none", "worse"]:
                self.Lup()
                if self.L == 1e9:
                    break
                direction = "worse"
            else:
                break
            if (best[1] - init_chi2) / init_chi2 < -0.1:
                break
        if nostep:
            if scarry_best[0] is not None:
                return scarry_best
            raise OptimizeStop("Could not find step to improve chi^2")
        return best

    @torch.no_grad()
    def fit(self):
        if len(self.current_state) == 0:
            if self.verbose > 0:
                AP_config.ap_logger.warning("No parameters to optimize. Exiting fit")
            return self
        self._covariance_matrix = None
        self.loss_history = [self._chi2(self.forward(self.current_state)).item()]
        self.L_history = [self.L]
        self.lambda_history = [self.current_state.detach().clone().cpu().numpy()]
        for iteration in range(self.max_iter):
            if self.verbose > 0:
                AP_config.ap_logger.info(f"Chi^2/DoF: {self.loss_history[-1]}, L: {self.L}")
            try:
                res = self.step(chi2=self.loss_history[-1])
            except OptimizeStop:
                if self.verbose > 0:
                    AP_config.ap_logger.warning("Could not find step to improve Chi^2, stopping")
                self.message = self.message + "fail. Could not find step to improve Chi^2"
                break
            self.L = res[2]
            self.current_state = (self.current_state + res[0]).detach()
            self.L_history.append(self.L)
            self.loss_history.append(res[1])
            self.lambda_history.append(self.current_state.detach().clone().cpu().numpy())
            self.Ldn()
            if len(self.loss_history) >= 3:
                if (self.loss_history[-3] - self.loss_history[-1]) / self.loss_history[-1] < \
                        self.relative_tolerance and self.L < 0.1:
                    self.message = self.message + "success"
                    break
            if len(self.loss_history) > 10:
                if (self.loss_history[-10] - self.loss_history[-1]) / self.loss_history[-1] < \
                        self.relative_tolerance:
                    self.message = self.message + "success by immobility. Convergence not guaranteed"
                    break
        else:
            self.message = self.message + "fail. Maximum iterations"
        if self.verbose > 0:
            AP_config.ap_logger.info(f"Final Chi^2/DoF: {self.loss_history[-1]}, L: {self.L_history[-1]}. "
                                     f"Converged: {self.message}")
        self.model.parameters.vector_set_representation(torch.as_tensor(self.res(), dtype=AP_config.ap_dtype))
        return self

    @torch.no_grad()
    def update_hess_grad(self, natural=False):
        """lm.py:408-425."""
        if self._sparse is not None:
            # J with respect to values = J_rep / (dvalue / drep), column by column
            J = self.jacobian(self.current_state)
            Y = self.forward(self.current_state)
            hess, grad = self._normal_equations(J, Y)
            if natural:
                d = torch.as_tensor(self.model.parameters.vector_dval_drep(self.current_state),
                                    dtype=torch.float64, device=hess.device).reshape(-1)
                hess = hess / d[:, None] / d[None, :]
                grad = grad / d
            self.hess, self.grad = hess, grad
            return
        if natural:
            vals = self.model.parameters.vector_transform_rep_to_val(self.current_state)
            J = self.model.jacobian(parameters=vals, as_representation=False).flatten("data").contiguous()
        else:
            J = self.jacobian(self.current_state)
        Y = self.forward(self.current_state)
        self.hess, self.grad = self._normal_equations(J, Y)

    @property
    def covariance_matrix(self):
        """lm.py:495-519."""
        if self._covariance_matrix is not None:
            return self._covariance_matrix
        self.update_hess_grad(natural=True)
        try:
            self._covariance_matrix = torch.linalg.inv(self.hess.cpu())
        except Exception:
            AP_config.ap_logger.warning("WARNING: Hessian is singular, likely at least one model is non-physical.")
            hess = self.hess.cpu()
            hess = hess + torch.eye(len(hess), dtype=hess.dtype) * (torch.diag(hess) == 0)
            self._covariance_matrix = torch.linalg.inv(hess)
        return self._covariance_matrix

    def update_uncertainty(self):
        """lm.py:521-539."""
        cov = self.covariance_matrix
        if torch.all(torch.isfinite(cov)):
            self.model.parameters.vector_set_uncertainty(torch.sqrt(torch.abs(torch.diag(cov))))
        else:
            AP_config.ap_logger.warning("Unable to update uncertainty due to non finite covariance matrix")
