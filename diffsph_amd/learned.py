"""Learned warm start of the DFSPH divergence solve.

The reference warm-starts every pressure solve with 0.75 p_prev
(src/diffSPH/v2/modules/dfsph.py:70-75).  Here a per-particle MLP
(16 -> 64 -> 64 -> 64 -> 1, bf16 weights and layer inputs, fp32 accumulation
on the MFMA units) maps neighbor-aggregated features of the previous
pressures and the current divergence source to the change of the
divergence-solve pressure; the initial guess is max(p_prev + prediction, 0).

* Training data are the solver's own converged pressure fields: a DFSPH run
  records (features before the solve, converged - previous pressure) per
  step (:func:`collect`).
* Training is on-policy (:func:`train_warm_start`): the first trajectory is
  warm-started with p_prev (the network's guess at zero output), later ones
  with the network itself, and the data are aggregated, so the network
  learns on the pressure states it produces.
* Training runs on the GPU with the library's fused forward+backward MFMA
  kernel and an Adam kernel (:meth:`WarmStartNet.train`).
* At run time (:meth:`WarmStartNet.install`) plain Jacobi is the fallback
  and the judge: a few Jacobi iterations from the network's guess and from
  the plain warm start decide which one starts the solve, and every solve
  runs to the solver's own convergence criterion (a bad network can cost
  iterations, never accuracy).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib, ptr
from .dfsph import JudgedGuess
from .neighborhood import WORKSPACE

NFEAT = 8       # used feature columns


def _pair_w(solver):
    """The density pass's W_ij of the current list, when the solver keeps it."""
    return solver.wpair if getattr(solver, "keep_pair_w", False) else None


def _records(solver, name: str = "_feat_rec"):
    """Grow-only [n][4] scratch for the packed neighbor fields of the feature
    pass (used with the kept W_ij), kept on the solver under `name`."""
    if _pair_w(solver) is None:
        return None
    r = getattr(solver, name, None)
    if r is None or r.shape[0] < solver.n:
        r = torch.empty(solver.n, 4, dtype=torch.float32, device=solver.pos.device)
        setattr(solver, name, r)
    return r


NIN = 16        # padded input width of the network


@dataclass
class JudgeStats:
    judged: int = 0
    used_net: int = 0
    used_plain: int = 0
    err_net: float = 0.0
    err_plain: float = 0.0


class WarmStartNet:
    def __init__(self, device, seed: int = 0, rings: int = 1):
        """`rings` 2 adds three second-ring features (columns 8-10: the
        kernel-weighted neighbor sums of the first ring's sums of
        p_div_prev, s and p_den_prev) from one more pass over the list
        (`dsph_ws_features_ring`).  At 1M particles it takes the learned
        divergence solve from 9.5 to 8.2 iterations per step."""
        L = lib()
        if rings not in (1, 2):
            raise ValueError("rings must be 1 or 2")
        self.rings = rings
        self.nfeat = NFEAT + 3 * (rings - 1)
        self.device = torch.device(device)
        self._raw = None  # (zeros, ones) [NIN]: the identity normalisation
        self.nparam = L.dsph_mlp_num_params()
        assert L.dsph_mlp_num_inputs() == NIN
        g = torch.Generator().manual_seed(seed)
        p = torch.zeros(self.nparam)
        # He initialisation of the weights, zero biases (flat layout of the C ABI)
        offs = self._offsets()
        for name, (o, shape) in offs.items():
            if name.startswith("W") or name == "w4":
                fan_in = shape[0]
                p[o:o + math.prod(shape)] = torch.randn(math.prod(shape), generator=g) * math.sqrt(2.0 / fan_in)
        # the output layer starts near zero: the initial guess is p_prev
        o, shape = offs["w4"]
        p[o:o + 64] *= 0.01
        self.params = p.to(self.device)
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.packed = torch.empty(L.dsph_mlp_packed_bytes(), dtype=torch.uint8, device=self.device)
        self.mean = torch.zeros(NIN, device=self.device)
        self.istd = torch.ones(NIN, device=self.device)
        self.t_mean = 0.0
        self.t_std = 1.0
        self.adam_step = 0
        self.judge = JudgeStats()
        self.pack()

    @staticmethod
    def _offsets():
        o = {}
        k = 0
        for name, shape in [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)),
                            ("b3", (64,)), ("w4", (64,)), ("b4", (1,))]:
            o[name] = (k, shape)
            k += math.prod(shape)
        return o

    def pack(self):
        check(lib().dsph_mlp_pack(ptr(self.params), ptr(self.packed), _lib.stream_handle(self.device)), "mlp_pack")

    # -- inference --------------------------------------------------------
    def forward(self, X: torch.Tensor) -> torch.Tensor:
        n = X.shape[0]
        y = torch.empty(n, device=X.device)
        check(lib().dsph_mlp_forward(ptr(X), n, ptr(self.params), ptr(self.packed), ptr(y),
                                     _lib.stream_handle(X.device)), "mlp_forward")
        return y

    def _feature_pass(self, solver, pdp, src, pnp, mean, istd) -> torch.Tensor:
        X = torch.empty(solver.n, NIN, device=self.device)
        pr = solver.nl.pairs()
        check(lib().dsph_ws_features(C.byref(pr), ptr(solver.vol), ptr(solver.rho), ptr(pdp), ptr(pnp),
                                     ptr(src), ptr(solver.alpha_div), C.c_float(solver.cfg.rho0),
                                     ptr(_pair_w(solver)), ptr(mean), ptr(istd), ptr(X), ptr(_records(solver)),
                                     _lib.stream_handle(self.device)), "ws_features")
        return X

    def features(self, solver, normalized: bool = True) -> torch.Tensor:
        n = solver.n
        pdp = solver.p_div if solver.p_div is not None else torch.zeros(n, device=self.device)
        pnp = solver.p_den if solver.p_den is not None else torch.zeros(n, device=self.device)
        if self._raw is None:
            self._raw = (torch.zeros(NIN, device=self.device), torch.ones(NIN, device=self.device))
        raw0, raw1 = self._raw
        if self.rings == 1:
            mean = self.mean if normalized else raw0
            istd = self.istd if normalized else raw1
            return self._feature_pass(solver, pdp, solver.div_src, pnp, mean, istd)
        mean = self.mean if normalized else raw0
        istd = self.istd if normalized else raw1
        pw, rec = _pair_w(solver), _records(solver)
        if pw is not None and rec is not None:
            X = torch.empty(n, NIN, device=self.device)
            ring = _records(solver, "_ring_rec")
            check(lib().dsph_ws_features_ring(C.byref(solver.nl.pairs()), ptr(solver.vol), ptr(solver.rho),
                                              ptr(pdp), ptr(pnp), ptr(solver.div_src), ptr(solver.alpha_div),
                                              C.c_float(solver.cfg.rho0), ptr(pw), ptr(mean), ptr(istd),
                                              self.nfeat, ptr(X), ptr(rec), ptr(ring),
                                              _lib.stream_handle(self.device)), "ws_features_ring")
            return X
        # without the kept W_ij: the first-ring pass over the first ring's sums
        X = self._feature_pass(solver, pdp, solver.div_src, pnp, raw0, raw1)
        X2 = self._feature_pass(solver, X[:, 1].contiguous(), X[:, 3].contiguous(), X[:, 5].contiguous(), raw0, raw1)
        X[:, NFEAT:NFEAT + 3] = X2[:, [1, 3, 5]]
        X = (X - mean) * istd
        X[:, self.nfeat:] = 0.0
        return X.contiguous()

    def predict_pressure(self, solver) -> torch.Tensor:
        X = self.features(solver)
        # p_prev + (y * t_std + t_mean), clamped at 0: in the forward pass's store
        p = torch.empty_like(solver.p_div)
        check(lib().dsph_mlp_forward_affine(ptr(X), X.shape[0], ptr(self.params), ptr(self.packed), ptr(solver.p_div),
                                            float(self.t_std), float(self.t_mean),
                                            1 if solver.cfg.clamp_pressure else 0, ptr(p),
                                            _lib.stream_handle(X.device)), "mlp_forward_affine")
        return p

    # -- training -----------------------------------------------------------
    def fit_normalization(self, X_raw: torch.Tensor, T_raw: torch.Tensor):
        nf = self.nfeat
        m = X_raw[:, :nf].mean(0)
        s = X_raw[:, :nf].std(0)
        self.mean.zero_()
        self.istd.fill_(1.0)
        self.mean[:nf] = m
        self.istd[:nf] = 1.0 / torch.clamp(s, min=1e-12)
        self.t_mean = float(T_raw.mean())
        self.t_std = float(max(T_raw.std(), 1e-12))

    def normalize(self, X_raw, T_raw):
        X = (X_raw - self.mean) * self.istd
        X[:, self.nfeat:] = 0.0
        T = (T_raw - self.t_mean) / self.t_std
        return X.contiguous(), T.contiguous()

    def loss_grad(self, X: torch.Tensor, T: torch.Tensor):
        L = lib()
        n = X.shape[0]
        ws = WORKSPACE.get(L.dsph_mlp_train_workspace(n), X.device)
        loss = torch.empty(1, device=X.device)
        check(L.dsph_mlp_loss_grad(ptr(X), ptr(T), n, C.c_float(1.0 / max(n, 1)), ptr(self.params),
                                   ptr(self.packed), ptr(self.grad), ptr(loss), ptr(ws), ws.numel(),
                                   _lib.stream_handle(X.device)), "mlp_loss_grad")
        return loss

    def train(self, X: torch.Tensor, T: torch.Tensor, steps: int = 300, lr: float = 2e-3, batch: int = 1 << 19,
              weight_decay: float = 0.0, seed: int = 0, log_every: int = 0):
        """Adam on minibatches of the normalised samples; returns loss history."""
        L = lib()
        n = X.shape[0]
        g = torch.Generator(device=X.device).manual_seed(seed)
        hist = []
        for s in range(steps):
            if n > batch:
                idx = torch.randint(0, n, (batch,), device=X.device, generator=g)
                xb, tb = X[idx].contiguous(), T[idx].contiguous()
            else:
                xb, tb = X, T
            loss = self.loss_grad(xb, tb)
            self.adam_step += 1
            lr_s = lr * 0.5 * (1.0 + math.cos(math.pi * s / max(steps, 1)))  # cosine decay
            check(L.dsph_adam_step(ptr(self.params), ptr(self.grad), ptr(self.m1), ptr(self.m2), self.nparam,
                                   C.c_float(lr_s), C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8),
                                   C.c_float(weight_decay), self.adam_step, _lib.stream_handle(X.device)), "adam")
            self.pack()
            if log_every and (s % log_every == 0 or s == steps - 1):
                hist.append((s, float(loss)))
        return hist

    def evaluate(self, X, T) -> float:
        """Mean squared error on normalised targets."""
        y = self.forward(X)
        return float(((y - T) ** 2).mean())

    # -- run time -----------------------------------------------------------
    def install(self, solver, judge_every: int = 8, probe_iters: int = 4, tolerance: float = 10.0,
                spike: float = 2.0):
        """Use the network as the divergence-solve warm start of `solver`.

        Plain Jacobi is the fallback and the judge.  On judged steps
        the solve's first `probe_iters` Jacobi iterations run from the
        network's guess, and as many from the solver's own warm start
        (warm_start * p_prev) on a copy; the network's guess is kept (the
        solve goes on) unless its error metric after those iterations
        exceeds `tolerance` times the plain one (floored at the convergence
        threshold) or is not finite (the solve starts again from the plain
        guess).  A guess that converges within them is kept.  The
        judge is deliberately lenient, a guard against a broken network
        rather than a per-step optimiser: the network's gain is over the
        trajectory (it keeps the low-frequency pressure that 0.75 p_prev
        throws away), while one step taken alone often favours the plain
        guess by a few iterations and switching to it loses that pressure.  Judged steps are the first one, every
        `judge_every`-th one (0: only forced ones) and any step after a
        solve that did not converge or took more than `spike` times the
        running mean of iterations.  Between judgements the last verdict is
        used.  Every solve still runs to the solver's own convergence
        criterion, so a bad network costs iterations, never accuracy."""
        state = {"k": 0, "use_net": True, "force": True, "mean_iters": None}
        thr = solver.cfg.threshold_div / solver.cfg.rho0

        def div_init(sv):
            if sv.p_div is None:
                return torch.zeros_like(sv.rho)
            judged = state["force"] or (judge_every > 0 and state["k"] % judge_every == 0)
            state["k"] += 1

            def plain():
                return sv.cfg.warm_start * sv.p_div

            if not judged and not state["use_net"]:
                self.judge.used_plain += 1
                return plain()
            p_net = self.predict_pressure(sv)
            if judged:
                self.judge.judged += 1
                state["force"] = False
                state["use_net"] = True
                self.judge.err_net, self.judge.err_plain = 0.0, float("nan")

                def verdict(sv, e_net):
                    # the network's error is the solve's own after probe_iters
                    # iterations; the plain guess's is probed on a copy
                    q = plain()
                    e_plain = sv.probe(q, probe_iters)[0]
                    ok = math.isfinite(e_net) and e_net <= tolerance * max(e_plain, thr)
                    self.judge.err_net, self.judge.err_plain = e_net, e_plain
                    state["use_net"] = ok
                    if ok:
                        return None
                    self.judge.used_net -= 1
                    self.judge.used_plain += 1
                    return q

                self.judge.used_net += 1
                return JudgedGuess(p_net, probe_iters, verdict)
            self.judge.used_net += 1
            return p_net

        def after_div(sv):
            it = sv._last_iters.get(False, 0)
            m = state["mean_iters"]
            if it >= sv.cfg.max_iters or (m is not None and it > spike * m + 2):
                state["force"] = True
            state["mean_iters"] = it if m is None else 0.8 * m + 0.2 * it

        solver.hooks["div_init"] = div_init
        solver.hooks["after_div"] = after_div
        solver.keep_pair_w = True
        return state


def collect(solver, steps: int, sample: float = 0.25, seed: int = 0, guess=None, rings: int = 1):
    """Run `steps` DFSPH steps and record (raw features, target) where
    target = converged divergence pressure - previous divergence pressure.
    `guess(solver)` gives the initial divergence pressure (default: the
    solver's own warm start, warm_start * p_prev).
    Returns (X_raw [N,16], T_raw [N]) on the solver's device."""
    net_for_features = WarmStartNet(solver.pos.device, rings=rings)
    Xs, Ts = [], []
    g = torch.Generator(device=solver.pos.device).manual_seed(seed)
    rec = {}

    def div_init(sv):
        if sv.p_div is None:
            rec.clear()
            return torch.zeros_like(sv.rho)
        rec["X"] = net_for_features.features(sv, normalized=False)
        rec["prev"] = sv.p_div.clone()
        return guess(sv) if guess is not None else sv.cfg.warm_start * sv.p_div

    def after_div(sv):
        if "X" not in rec:
            return
        n = sv.n
        k = max(1, int(n * sample))
        idx = torch.randperm(n, device=sv.pos.device, generator=g)[:k]
        Xs.append(rec["X"][idx])
        Ts.append((sv.p_div - rec["prev"])[idx])
        rec.clear()

    saved = dict(solver.hooks)
    solver.hooks["div_init"] = div_init
    solver.hooks["after_div"] = after_div
    solver.keep_pair_w = True
    try:
        for _ in range(steps):
            solver.step()
    finally:
        solver.hooks.clear()
        solver.hooks.update(saved)
    return torch.cat(Xs), torch.cat(Ts)


def train_warm_start(make_solver, rounds: int = 3, steps: int = 24, sample: float = 0.1, train_steps: int = 400,
                     lr: float = 2e-3, seed: int = 0, device=None, log=None, rings: int = 1) -> WarmStartNet:
    """Train a warm-start network on-policy.

    Round 0 records a trajectory warm-started with p_prev (the network's
    guess at zero output); every later round records a trajectory driven by
    the current network, adds it to the data and continues training, so the
    network learns on the pressure states it produces itself.
    `make_solver(r)` returns a fresh solver for round r."""
    net = None
    Xa, Ta = [], []
    for r in range(rounds):
        sv = make_solver(r)
        if net is None:
            net = WarmStartNet(device or sv.pos.device, seed=seed, rings=rings)
            X, T = collect(sv, steps, sample, seed=seed + r, guess=lambda s: s.p_div, rings=rings)
            net.fit_normalization(X, T)
        else:
            X, T = collect(sv, steps, sample, seed=seed + r, guess=net.predict_pressure, rings=rings)
        Xa.append(X)
        Ta.append(T)
        Xn, Tn = net.normalize(torch.cat(Xa), torch.cat(Ta))
        net.train(Xn, Tn, steps=train_steps, lr=lr, seed=seed + r)
        if log is not None:
            log(f"round {r}: {Xn.shape[0]} samples, normalized mse {net.evaluate(Xn, Tn):.4f}")
    return net


# ---------------------------------------------------------------------------
# Learned preconditioner: a residual correction inside the divergence solve
# ---------------------------------------------------------------------------
class CorrectionNet(WarmStartNet):
    """Learned preconditioner of the DFSPH divergence solve.

    After ``k1`` Jacobi iterations the remaining error is smooth (Jacobi
    removes the rough part first); a second per-particle MLP of the same
    shape and kernels maps neighbor-aggregated features of the current
    pressure p_k, the current residual r_k and the previous step's converged
    pressure to the remaining change p* - p_k, i.e. it applies a learned
    approximate inverse of the pressure operator to the residual.  Features
    (same kernel as the warm start): 0 p_k, 1 sum V W p_k, 2 r_k,
    3 sum V W r_k, 4 p_prev, 5 sum V W p_prev, 6 rho/rho0 - 1, 7 alpha."""

    def __init__(self, device, k1: int = 4, seed: int = 1):
        super().__init__(device, seed=seed)
        self.k1 = k1

    def features_at(self, solver, p: torch.Tensor, normalized: bool = True) -> torch.Tensor:
        n = solver.n
        X = torch.empty(n, NIN, device=self.device)
        prev = solver.p_div if solver.p_div is not None else torch.zeros(n, device=self.device)
        mean = self.mean if normalized else torch.zeros(NIN, device=self.device)
        istd = self.istd if normalized else torch.ones(NIN, device=self.device)
        pr = solver.nl.pairs()
        check(lib().dsph_ws_features(C.byref(pr), ptr(solver.vol), ptr(solver.rho), ptr(p), ptr(prev),
                                     ptr(solver.residual), ptr(solver.alpha_div), C.c_float(solver.cfg.rho0),
                                     ptr(_pair_w(solver)), ptr(mean), ptr(istd), ptr(X), ptr(_records(solver)),
                                     _lib.stream_handle(self.device)), "ws_features")
        return X

    def corrected(self, solver, p: torch.Tensor) -> torch.Tensor:
        y = self.forward(self.features_at(solver, p))
        q = p + (y * self.t_std + self.t_mean)
        return q.clamp_min_(0.0) if solver.cfg.clamp_pressure else q

    def install(self, solver, judge_every: int = 8, probe_iters: int = 2):
        """Apply the correction after k1 divergence iterations.  Plain
        Jacobi is the judge: on judged steps (the first, every
        `judge_every`-th) `probe_iters` iterations from the corrected and
        from the uncorrected pressure are compared and the correction is
        kept only if its error metric is not larger; the verdict holds until
        the next judgement.  The solve always ends on the solver's own
        stopping rule."""
        state = {"k": 0, "use": True}

        def fn(sv, p):
            q = self.corrected(sv, p)
            if judge_every > 0 and state["k"] % judge_every == 0:
                e_c = sv.probe_guess(q, probe_iters)[0]
                e_p = sv.probe_guess(p, probe_iters)[0]
                state["use"] = math.isfinite(e_c) and e_c <= e_p
                self.judge.judged += 1
                self.judge.err_net, self.judge.err_plain = e_c, e_p
            state["k"] += 1
            if state["use"]:
                self.judge.used_net += 1
                return q
            self.judge.used_plain += 1
            return p

        solver.hooks["div_split"] = (self.k1, fn)
        solver.keep_pair_w = True
        return state


def collect_correction(solver, steps: int, k1: int, sample: float = 0.1, seed: int = 0):
    """Run `steps` DFSPH steps recording, for every divergence solve that is
    not converged after k1 iterations, (raw features at k1, p* - p_k1)."""
    feat = CorrectionNet(solver.pos.device, k1=k1)
    Xs, Ts = [], []
    g = torch.Generator(device=solver.pos.device).manual_seed(seed)
    rec = {}

    def fn(sv, p):
        rec["X"] = feat.features_at(sv, p, normalized=False)
        rec["p"] = p.clone()
        return installed[1](sv, p) if installed is not None else p

    def after_div(sv):
        if "X" in rec:
            n = sv.n
            k = max(1, int(n * sample))
            idx = torch.randperm(n, device=sv.pos.device, generator=g)[:k]
            Xs.append(rec["X"][idx])
            Ts.append((sv.p_div - rec["p"])[idx])
        rec.clear()
        if chained is not None:
            chained(sv)

    chained = solver.hooks.get("after_div")
    installed = solver.hooks.get("div_split")
    saved = dict(solver.hooks)
    solver.hooks["div_split"] = (k1, fn)
    solver.hooks["after_div"] = after_div
    solver.keep_pair_w = True
    try:
        for _ in range(steps):
            solver.step()
    finally:
        solver.hooks.clear()
        solver.hooks.update(saved)
    if not Xs:
        return None, None
    return torch.cat(Xs), torch.cat(Ts)


def train_correction(make_solver, warm: WarmStartNet | None = None, k1: int = 4, rounds: int = 2, steps: int = 24,
                     sample: float = 0.1, train_steps: int = 400, lr: float = 2e-3, seed: int = 0,
                     log=None) -> CorrectionNet:
    """Train the residual-correction network on the solver's own solves
    (warm-started by `warm` when given, as it will run).  Round 0 records
    uncorrected solves; later rounds run with the current network installed
    (on-policy) and add their samples."""
    net = None
    Xa, Ta = [], []
    for r in range(rounds):
        sv = make_solver(r)
        if warm is not None:
            warm.install(sv)
        if net is not None:
            net.install(sv, judge_every=0)
        X, T = collect_correction(sv, steps, k1, sample, seed=seed + r)
        if X is None:
            break
        if net is None:
            net = CorrectionNet(sv.pos.device, k1=k1, seed=seed + 1)
            net.fit_normalization(X, T)
        Xa.append(X)
        Ta.append(T)
        Xn, Tn = net.normalize(torch.cat(Xa), torch.cat(Ta))
        net.train(Xn, Tn, steps=train_steps, lr=lr, seed=seed + r)
        if log is not None:
            log(f"correction round {r}: {Xn.shape[0]} samples, normalized mse {net.evaluate(Xn, Tn):.4f}")
    return net
