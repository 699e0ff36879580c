"""Differentiable SOAP model: FP64 forward and adjoint on the GPU (soap_diff_* of
the C ABI), exposed as a torch.autograd.Function.

The model (no counterpart in the reference; built from its shift, convolution,
mask CCF and Gaussian fit): disk cells with rigid rotation delta = vsini y and
quadratic limb darkening; per-phase spot and facula filling-factor maps (M x M
over [-1, 1]^2, bilinear) blending the quiet seed rows Q, the spot rows S scaled
by the spot contrast and, optionally, the facula rows Fa scaled by the facula
contrast and the reference's facula intensity factor 1 + LB(mu)
(SOAP_integration.cu:317-323); F_p = instrumental convolution of the phase
spectrum; RV and FWHM from the Gaussian fit of the raw mask CCF of F_p, polished
to the least-squares optimum; BIS = the bisector span of that CCF at the polished Gaussian
(BIS_FIT2.cpp:214-268).  Loss = w_spec sum (F - T)^2 (trimmed pixels)
+ w_rv sum (RV - RV_t)^2 + w_fwhm sum (FWHM - FWHM_t)^2 + w_bis sum (BIS - BIS_t)^2.
Gradients: vsini, u1, u2, both contrasts, the seed rows Q, S, Fa and every pixel of
both maps.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .engine import GRID_LOG, DiffParams, Engine, SoapError, _d

_dp = ctypes.POINTER(ctypes.c_double)


def _dev(t: torch.Tensor) -> ctypes.c_void_p:
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise SoapError("expected a contiguous float64 tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


class DiffModel:
    """One differentiable problem on an Engine: grid, mu bins, mask and sizes fixed."""

    def __init__(self, eng: Engine, wave, mu_bounds, line_l, line_w, vgrid, grid: int, n_phase: int,
                 map_res: int, sigma_ms: float, trim: int = 210, lb=None):
        """lb: the facula intensity polynomial LB(mu) (4 values, constant first), default zeros."""
        self.eng = eng
        eng.set_wave(wave, GRID_LOG)
        eng.set_mu_bounds(mu_bounds)
        eng.set_mask(line_l, line_w, vgrid, trim)
        self.lb = np.zeros(4) if lb is None else np.asarray(lb, dtype=np.float64).reshape(4)
        lbv, lbp = _d(self.lb)
        eng._chk(eng.L.soap_diff_setup(eng.ctx, int(grid), int(n_phase), int(map_res), float(sigma_ms), lbp),
                 "diff_setup")
        self.n = len(wave)
        self.n_mu = len(mu_bounds) + 1
        self.P = int(n_phase)
        self.M = int(map_res)
        self.n_v = len(vgrid)
        self._fwd_id = 0
        self._keep = None

    def forward(self, params, q, s, maps, target, rv_t, fw_t, w_spec=1.0, w_rv=0.0, w_fw=0.0, bis_t=None,
                w_bis=0.0, f=None, fmaps=None):
        """params = (vsini [km/s], u1, u2, contrast[, contrast_fac]); q, s [n_mu, n], maps [P, M, M],
        target [P, n] float64 GPU tensors; f [n_mu, n] and fmaps [P, M, M] (together, optional): the
        facula population.  Returns (loss, rv[P], fwhm[P], bis[P])."""
        pv = [float(x) for x in params]
        vs, u1, u2, kap = pv[:4]
        kf = pv[4] if len(pv) > 4 else 0.0
        if (f is None) != (fmaps is None):
            raise SoapError("facula seeds and facula maps go together")
        torch.cuda.synchronize()
        prm = DiffParams(vs, u1, u2, kap, kf)
        rt, rtp = _d(np.asarray(rv_t, dtype=np.float64).reshape(self.P))
        ft, ftp = _d(np.asarray(fw_t, dtype=np.float64).reshape(self.P))
        bt, btp = _d(np.zeros(self.P) if bis_t is None else np.asarray(bis_t, dtype=np.float64).reshape(self.P))
        loss = ctypes.c_double()
        rv, rvp = _d(np.zeros(self.P))
        fw, fwp = _d(np.zeros(self.P))
        bs, bsp = _d(np.zeros(self.P))
        self._keep = (q, s, maps, target, f, fmaps)  # the engine reads them again in backward
        nul = ctypes.c_void_p()
        self.eng._chk(self.eng.L.soap_diff_forward(self.eng.ctx, ctypes.byref(prm), _dev(q), _dev(s),
                                                   nul if f is None else _dev(f), _dev(maps),
                                                   nul if fmaps is None else _dev(fmaps), _dev(target), rtp, ftp,
                                                   btp, float(w_spec), float(w_rv), float(w_fw), float(w_bis),
                                                   ctypes.byref(loss), rvp, fwp, bsp),
                      "diff_forward")
        self._fwd_id += 1
        return loss.value, rv, fw, bs

    def backward(self):
        """Gradients of the last forward's loss: (g5 = d/d(vsini, u1, u2, contrast, contrast_fac), gQ, gS,
        gmaps, gFa, gfmaps); gFa and gfmaps are None when the forward had no faculae."""
        q = self._keep[0]
        fac = self._keep[4] is not None
        gq = torch.empty((self.n_mu, self.n), dtype=torch.float64, device=q.device)
        gs = torch.empty_like(gq)
        gm = torch.empty((self.P, self.M, self.M), dtype=torch.float64, device=q.device)
        gf = torch.empty_like(gq) if fac else None
        gfm = torch.empty_like(gm) if fac else None
        g5, g5p = _d(np.zeros(5))
        torch.cuda.synchronize()
        nul = ctypes.c_void_p()
        self.eng._chk(self.eng.L.soap_diff_backward(self.eng.ctx, g5p, _dev(gq), _dev(gs),
                                                    _dev(gf) if fac else nul, _dev(gm), _dev(gfm) if fac else nul),
                      "diff_backward")
        return g5, gq, gs, gm, gf, gfm

    def get(self, what: str):
        """Host copies of the last forward: 'F' [P, n], 'ccf' [P, n_v], 'theta' [P, 4] = (c, k, RV, FWHM),
        'bis' [P] (bisector spans)."""
        shape = {"F": (self.P, self.n), "ccf": (self.P, self.n_v), "theta": (self.P, 4), "bis": (self.P,)}[what]
        out = np.empty(shape)
        self.eng._chk(self.eng.L.soap_diff_get(self.eng.ctx, ("F", "ccf", "theta", "bis").index(what),
                                               out.ctypes.data_as(_dp)), "diff_get")
        return out


class SoapLoss(torch.autograd.Function):
    """loss = SoapLoss.apply(model, params, q, s, maps, target, rv_t, fw_t, weights, bis_t=None, f=None,
    fmaps=None) with weights = (w_spec, w_rv, w_fw) or (w_spec, w_rv, w_fw, w_bis).  params is a
    float64 tensor (vsini, u1, u2, contrast) or (..., contrast_fac); f and fmaps the optional
    facula seed rows and maps.  Gradients flow to params, q, s, maps, f and fmaps.  One
    backward per forward (the engine keeps the last forward)."""

    @staticmethod
    def forward(ctx, model, params, q, s, maps, target, rv_t, fw_t, weights, bis_t=None, f=None, fmaps=None):
        w = tuple(weights) + (0.0,) * (4 - len(weights))
        loss, rv, fw, bs = model.forward(params.detach().cpu().numpy(), q.detach(), s.detach(), maps.detach(),
                                         target.detach(), rv_t, fw_t, w[0], w[1], w[2], bis_t, w[3],
                                         f=None if f is None else f.detach(),
                                         fmaps=None if fmaps is None else fmaps.detach())
        ctx.model = model
        ctx.fwd_id = model._fwd_id
        ctx.params_device = params.device
        ctx.n_params = params.numel()
        ctx.rv, ctx.fw, ctx.bis = rv, fw, bs
        return torch.tensor(loss, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, gout):
        if ctx.model._fwd_id != ctx.fwd_id:
            raise SoapError("SoapLoss: the model ran another forward since this graph was built")
        g5, gq, gs, gm, gf, gfm = ctx.model.backward()
        gp = torch.as_tensor(g5[:ctx.n_params], dtype=torch.float64, device=ctx.params_device)
        gp = gp * gout.to(ctx.params_device)
        go = gout.to(gq.device)
        return (None, gp, gq * go, gs * go, gm * go, None, None, None, None, None,
                None if gf is None else gf * go, None if gfm is None else gfm * go)


class DiffBatch:
    """Forward + adjoint of many DiffModels as one HIP graph (soap_diff_batch_* of
    the C ABI).  Every input, parameter and result is a device tensor bound once;
    a step replays the graph on the values those tensors hold."""

    def __init__(self, models):
        from .engine import load
        self.L = load()
        self.models = list(models)
        n = len(self.models)
        ctxs = (ctypes.c_void_p * n)(*[m.eng.ctx for m in self.models])
        self.b = ctypes.c_void_p()
        rc = self.L.soap_diff_batch_create(ctxs, n, ctypes.byref(self.b))
        if rc != 0:
            msg = self.L.soap_diff_batch_last_error(self.b) if self.b else b""
            self.close()
            raise SoapError(f"soap_diff_batch_create failed ({rc}): {msg.decode() if msg else ''}")
        self._keep = [None] * n

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.soap_diff_batch_last_error(self.b)
            raise SoapError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def bind(self, i, vsini_max, params, q, s, maps, target, rv_t, fw_t, bis_t, weights, loss, g5, gq, gs, gm,
             f=None, fmaps=None, gf=None, gfm=None):
        """All tensors float64, contiguous, on the GPU: params [5] = (vsini, u1, u2, contrast,
        contrast_fac), q / s [n_mu][n], maps [P][M][M], target [P][n], rv_t / fw_t / bis_t [P];
        weights (w_spec, w_rv, w_fw, w_bis); outputs loss [1], g5 [5], gq / gs like q, gm like
        maps; the facula population (all four or none): f like q, fmaps like maps and their
        gradients gf, gfm."""
        fac = (f, fmaps, gf, gfm)
        if any(t is None for t in fac) and not all(t is None for t in fac):
            raise SoapError("facula seeds, maps and their gradients go together")
        nul = ctypes.c_void_p()
        dv = lambda t: nul if t is None else _dev(t)  # noqa: E731
        w_spec, w_rv, w_fw, w_bis = (float(x) for x in tuple(weights) + (0.0,) * (4 - len(weights)))
        self._chk(self.L.soap_diff_batch_bind(self.b, int(i), float(vsini_max), _dev(params), _dev(q), _dev(s),
                                              dv(f), _dev(maps), dv(fmaps), _dev(target), _dev(rv_t), _dev(fw_t),
                                              _dev(bis_t), w_spec, w_rv, w_fw, w_bis, _dev(loss), _dev(g5), _dev(gq),
                                              _dev(gs), dv(gf), _dev(gm), dv(gfm)), "diff_batch_bind")
        self._keep[i] = (params, q, s, maps, target, rv_t, fw_t, bis_t, loss, g5, gq, gs, gm, f, fmaps, gf, gfm)

    def step(self, graph: bool = True, adjoint: bool = True):
        """Forward + loss (+ adjoint when adjoint) of every bound instance."""
        self._chk(self.L.soap_diff_batch_step(self.b, int(bool(graph)), int(bool(adjoint))), "diff_batch_step")

    def sync(self):
        self._chk(self.L.soap_diff_batch_sync(self.b), "diff_batch_sync")

    def close(self):
        if self.b:
            self.L.soap_diff_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
