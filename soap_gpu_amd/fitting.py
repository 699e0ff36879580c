"""Fitting loop over many differentiable instances: L-BFGS on the contrasts, vsini
and the filling-factor maps of each instance (spot map, and facula map when the
instance carries the facula population), against target RV, FWHM and BIS
(bisector span) series (and optionally target spectra), with the FP64 adjoint of
soap_diff_* for the gradients.

No counterpart in the reference.  Each instance is a DiffModel on its own
Engine (its own context and stream); the objective is the sum of the
instances' losses, so the gradient of each instance's parameters is that
instance's own gradient.  The maps are kept in [0, 1] through a logistic
parameterisation, maps = 1 / (1 + exp(-z)).  The forward of every instance is
recomputed at each evaluation; the shifted spectra are never stored between
forward and adjoint beyond one instance's buffers (recompute-in-backward).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .diff import DiffBatch, DiffModel, SoapLoss
from .engine import Engine, select_mask_lines


class Instance:
    """One fitting problem: a DiffModel with its seeds, targets and starting point."""

    def __init__(self, w, maps0, n_phase: int, map_res: int, target_spec=None, rv_t=None, fw_t=None, bis_t=None,
                 weights=(0.0, 1e-3, 1e-5, 1e-3), u=(0.29, 0.34), vsini0: float = 2.0, contrast0: float = 0.3,
                 trim: int = 210, device: int = 0, fmaps0=None, contrast_fac0: float = 1.1):
        """weights = (w_spec, w_rv, w_fwhm[, w_bis]); fmaps0 (optional): starting facula maps, which
        switch on the facula population (seed rows w.cubes["fac"], intensity 1 + LB(mu) with the
        star's w.star["lb"]) with the facula contrast starting at contrast_fac0."""
        self.eng = Engine(device)
        sl, sw = select_mask_lines(w.wave, w.mask_l, w.mask_w, trim)
        self.model = DiffModel(self.eng, w.wave, w.mu_bounds, sl, sw, w.vgrid, w.star["grid"], n_phase, map_res,
                               w.sigma_inst, trim, lb=w.star.get("lb"))
        dev = torch.device("cuda", device)
        self.q = torch.tensor(w.cubes["q"], dtype=torch.float64, device=dev)
        self.s = torch.tensor(w.cubes["sp"], dtype=torch.float64, device=dev)
        self.facula = fmaps0 is not None
        self.f = torch.tensor(w.cubes["fac"], dtype=torch.float64, device=dev) if self.facula else None
        self.P, self.M, self.n = n_phase, map_res, w.n
        self.target = (torch.zeros((n_phase, w.n), dtype=torch.float64, device=dev) if target_spec is None
                       else torch.as_tensor(target_spec, dtype=torch.float64, device=dev))
        self.rv_t = np.zeros(n_phase) if rv_t is None else np.asarray(rv_t, dtype=np.float64)
        self.fw_t = np.zeros(n_phase) if fw_t is None else np.asarray(fw_t, dtype=np.float64)
        self.bis_t = np.zeros(n_phase) if bis_t is None else np.asarray(bis_t, dtype=np.float64)
        self.weights = tuple(float(x) for x in weights) + (0.0,) * (4 - len(weights))
        self.u = tuple(float(x) for x in u)
        def logit(a):
            a = np.clip(np.asarray(a, dtype=np.float64), 1e-6, 1 - 1e-6)
            return torch.tensor(np.log(a / (1 - a)), dtype=torch.float64, device=dev, requires_grad=True)

        # fitted: (vsini, contrast[, contrast_fac]) and the map logits
        th0 = [vsini0, contrast0] + ([contrast_fac0] if self.facula else [])
        self.theta = torch.tensor(th0, dtype=torch.float64, device=dev, requires_grad=True)
        self._u = torch.tensor(self.u, dtype=torch.float64, device=dev)
        self.z = logit(maps0)
        self.zf = logit(fmaps0) if self.facula else None

    def maps(self):
        return torch.sigmoid(self.z)

    def fmaps(self):
        return torch.sigmoid(self.zf) if self.facula else None

    def variables(self):
        return [self.theta, self.z] + ([self.zf] if self.facula else [])

    def params(self):
        p = [self.theta[0], self._u[0], self._u[1], self.theta[1]]
        return torch.stack(p + ([self.theta[2]] if self.facula else []))

    def loss(self):
        return SoapLoss.apply(self.model, self.params(), self.q, self.s, self.maps(), self.target, self.rv_t,
                              self.fw_t, self.weights, self.bis_t, self.f, self.fmaps())

    def observe(self, params, maps, fmaps=None):
        """RV, FWHM and BIS series of the model at given parameters (targets of a synthetic problem)."""
        dev = self.q.device
        fm = None if fmaps is None else torch.as_tensor(fmaps, dtype=torch.float64, device=dev)
        _, rv, fw, bis = self.model.forward(params, self.q, self.s, torch.as_tensor(maps, dtype=torch.float64,
                                                                                    device=dev),
                                            self.target, self.rv_t, self.fw_t, 0.0, 0.0, 0.0,
                                            f=self.f if fm is not None else None, fmaps=fm)
        return rv, fw, bis

    def close(self):
        self.eng.close()


def fit(instances, iters: int = 50, history: int = 10, lr: float = 1.0, log=None):
    """L-BFGS (strong-Wolfe line search) on the summed loss of all instances.
    Returns dict(loss0, loss, iters, evals, seconds, ms_per_eval)."""
    params = [t for inst in instances for t in inst.variables()]
    opt = torch.optim.LBFGS(params, lr=lr, max_iter=iters, history_size=history, line_search_fn="strong_wolfe",
                            tolerance_grad=0.0, tolerance_change=0.0)
    evals = [0]
    hist = []

    def closure():
        opt.zero_grad()
        total = torch.zeros((), dtype=torch.float64)
        for inst in instances:
            l_i = inst.loss()
            l_i.backward()
            total = total + l_i.detach().cpu()
        evals[0] += 1
        hist.append(float(total))
        if log and evals[0] % 10 == 0:
            log(f"[fit] eval {evals[0]}: loss {float(total):.6e}")
        return total

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss0 = float(closure())
    opt.step(closure)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = opt.state[opt._params[0]]
    return dict(loss0=loss0, loss=hist[-1], best=min(hist), iters=int(st.get("n_iter", 0)), evals=evals[0],
                seconds=dt, ms_per_eval=dt / evals[0] * 1e3, history=hist)


def fit_graph(instances, iters: int = 50, history: int = 10, lr: float = 1.0, vsini_max=None, log=None):
    """As fit(), with each objective evaluation (forward + adjoint of every
    instance) one replay of a captured HIP graph (DiffBatch).  The fitted values of
    all instances live in two flat tensors (vsini and the contrasts per instance,
    and the concatenated map logits: per instance its spot maps, then its facula
    maps when the instances carry faculae) whose slices are bound to the graph, so
    an evaluation is a handful of tensor operations plus one graph launch; the
    results are written back to each instance's theta, z and zf.  Either every
    instance carries the facula population or none does.  vsini_max bounds the
    fitted vsini of each instance (default 2 x its start + 5 km/s); the loss is NaN
    beyond it."""
    dev = instances[0].q.device
    n = len(instances)
    fac = instances[0].facula
    if any(inst.facula != fac for inst in instances):
        raise ValueError("fit_graph: either every instance carries faculae or none does")
    nmap = 2 if fac else 1
    sizes = [inst.z.numel() * nmap for inst in instances]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    theta = torch.stack([inst.theta.detach() for inst in instances]).clone().requires_grad_(True)  # (n, 2 or 3)
    z = torch.cat([t.detach().reshape(-1) for inst in instances for t in ([inst.z, inst.zf] if fac else [inst.z])])
    z = z.clone().requires_grad_(True)
    u = torch.tensor([inst.u for inst in instances], dtype=torch.float64, device=dev)  # (n, 2)
    prm = torch.zeros((n, 5), dtype=torch.float64, device=dev)
    maps = torch.empty_like(z).detach()
    loss = torch.empty(n, dtype=torch.float64, device=dev)
    g5 = torch.empty((n, 5), dtype=torch.float64, device=dev)
    gm = torch.empty_like(maps)
    batch = DiffBatch([inst.model for inst in instances])
    keep = []
    for i, inst in enumerate(instances):
        vmax = float(vsini_max[i]) if vsini_max is not None else 2.0 * float(inst.theta[0].detach()) + 5.0
        shape = inst.z.shape
        m1 = int(offs[i]) + inst.z.numel()
        t = dict(rv_t=torch.as_tensor(inst.rv_t, dtype=torch.float64, device=dev).contiguous(),
                 fw_t=torch.as_tensor(inst.fw_t, dtype=torch.float64, device=dev).contiguous(),
                 bis_t=torch.as_tensor(inst.bis_t, dtype=torch.float64, device=dev).contiguous(),
                 gq=torch.empty_like(inst.q), gs=torch.empty_like(inst.s),
                 gf=torch.empty_like(inst.q) if fac else None)
        sl = slice(int(offs[i]), m1)
        slf = slice(m1, int(offs[i + 1]))
        batch.bind(i, vmax, prm[i], inst.q, inst.s, maps[sl].view(shape), inst.target, t["rv_t"], t["fw_t"],
                   t["bis_t"], inst.weights, loss[i:i + 1], g5[i], t["gq"], t["gs"], gm[sl].view(shape),
                   f=inst.f if fac else None, fmaps=maps[slf].view(shape) if fac else None, gf=t["gf"],
                   gfm=gm[slf].view(shape) if fac else None)
        keep.append(t)
    opt = torch.optim.LBFGS([theta, z], lr=lr, max_iter=iters, history_size=history, line_search_fn="strong_wolfe",
                            tolerance_grad=0.0, tolerance_change=0.0)
    evals = [0]
    hist = []
    cols = [0, 3, 4] if fac else [0, 3]

    def evaluate():
        with torch.no_grad():
            prm[:, 0] = theta[:, 0]
            prm[:, 1:3] = u
            prm[:, 3] = theta[:, 1]
            if fac:
                prm[:, 4] = theta[:, 2]
            m = torch.sigmoid(z)
            maps.copy_(m)
        torch.cuda.synchronize()
        batch.step(graph=True)
        batch.sync()
        with torch.no_grad():
            theta.grad = g5[:, cols].clone()
            z.grad = gm * m * (1 - m)
            return loss.sum()

    def closure():
        opt.zero_grad()
        total = evaluate()
        evals[0] += 1
        hist.append(float(total))
        if log and evals[0] % 10 == 0:
            log(f"[fit] eval {evals[0]}: loss {hist[-1]:.6e}")
        return total

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss0 = float(closure())
    opt.step(closure)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = opt.state[opt._params[0]]
    # the device part of one evaluation alone: graph replays at the final parameters,
    # forward + adjoint and the forward (with the loss) alone
    times = {}
    for adj in (True, False):
        batch.step(graph=True, adjoint=adj)
        batch.sync()
        t1 = time.perf_counter()
        for _ in range(3):
            batch.step(graph=True, adjoint=adj)
        batch.sync()
        times[adj] = (time.perf_counter() - t1) / 3 * 1e3
    batch.close()
    with torch.no_grad():
        for i, inst in enumerate(instances):
            inst.theta.copy_(theta[i])
            m1 = int(offs[i]) + inst.z.numel()
            inst.z.copy_(z[int(offs[i]):m1].view(inst.z.shape))
            if fac:
                inst.zf.copy_(z[m1:int(offs[i + 1])].view(inst.zf.shape))
    return dict(loss0=loss0, loss=hist[-1], best=min(h for h in hist if h == h), iters=int(st.get("n_iter", 0)),
                evals=evals[0], seconds=dt, ms_per_eval=dt / evals[0] * 1e3, graph_ms=times[True],
                graph_fwd_ms=times[False], history=hist)
