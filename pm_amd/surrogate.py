"""Learned VTI traveltime surrogate: data, training and evaluation.

The network (5 -> 64 -> 64 -> 64 -> 64 -> 2, SiLU) maps the encoded
(tau, s_s, s_r, v_rms, eta) of a (pillar, trace, t0) triple (tau = t0 / 2,
s = sine of a leg's straight-ray angle) to out_scale * (c_s, c_r), the
anisotropic factor of each one-way leg:

    t_leg = sqrt(tau^2 + (d / v)^2 (1 + c)),

where the exact leg time is the VTI one-way time of
pstm/algorithm/anisotropy_vti.py:181-220 (for which c = -2 eta s^2 /
(1 + 2 eta s^2); the network learns it from the exact times).  Training
samples are drawn on the job's own v_rms and eta fields (random x, y, t0 in
the model, random source and receiver distances) and their targets come from
an f64 HIP kernel; forward, backward and Adam are HIP kernels
(pm_amd/csrc/pstm_mlp.hip), and the migration evaluates the same network
inside its gather with BF16 MFMA (pm_amd/csrc/pstm_hip.hip).  Tensors are
device memory only.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from pm_amd.kernels import _abi


def _p(t) -> int:
    return t.data_ptr()


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(lib, rc: int, what: str) -> None:
    if rc != 0:
        msg = lib.pstm_mlp_last_error().decode()
        raise (ValueError if rc == 1 else RuntimeError)(f"{what}: {msg}")


@dataclass
class Normalizer:
    """Input encoding of the network.

    The five inputs are (tau, d_s, d_r, v, eta).  The distances enter through
    the sine of the straight-ray angle of each leg, s = d / (v t_hyp) with
    t_hyp = sqrt(tau^2 + d^2 / v^2) (s = 0 when t_hyp = 0): for the VTI time
    the leg factor is a function of (s, eta) alone,
    c = -2 eta s^2 / (1 + 2 eta s^2), while d / (v tau) itself is unbounded at
    shallow times.  Every input is then mapped affinely to
    about [-1, 1]."""

    tau_max: float
    d_max: float
    v_lo: float
    v_hi: float
    eta_max: float
    out_scale: float = 1.5    # network outputs are out_scale * the two legs' factors
                              # (VTI: g = c / eta in [-2, 0]; curved ray: c)

    def encode(self, tau, d_s, d_r, v, eta):
        """Raw inputs (torch, any float dtype) -> (n, 5) f32 network inputs."""
        import torch

        def sine(d):
            th = hyperbolic(tau, d, v)
            return torch.where(th > 0, d / (v * th.clamp_min(1e-300)), torch.zeros_like(th))

        cols = [2.0 * tau / self.tau_max - 1.0, 2.0 * sine(d_s) - 1.0, 2.0 * sine(d_r) - 1.0,
                2.0 * (v - self.v_lo) / (self.v_hi - self.v_lo) - 1.0,
                2.0 * eta / max(self.eta_max, 1e-6) - 1.0]
        return torch.stack(cols, 1).float().contiguous()

    def encode_curved(self, tau, d_s, d_r, vg, k):
        """Raw curved-ray inputs (tau, d_s, d_r, vg, k) -> (n, 5) f32 network
        inputs; the sines are k d / sqrt(vg^2 + k^2 d^2)."""
        import torch

        def sine(d):
            kd = k * d
            return kd / (vg * vg + kd * kd).sqrt()

        cols = [2.0 * tau / self.tau_max - 1.0, 2.0 * sine(d_s) - 1.0, 2.0 * sine(d_r) - 1.0,
                2.0 * (vg - self.v_lo) / (self.v_hi - self.v_lo) - 1.0,
                2.0 * k / max(self.eta_max, 1e-6) - 1.0]
        return torch.stack(cols, 1).float().contiguous()

    def apply(self, x):
        """x (n, 5) raw (tau, d_s, d_r, v, eta) -> encoded f32."""
        return self.encode(x[:, 0], x[:, 1], x[:, 2], x[:, 3], x[:, 4])


def init_params(seed: int = 0) -> np.ndarray:
    """Seeded He-uniform weights (SiLU ~ ReLU gain), zero biases, in the
    library's layout (per layer W (out x in) then b)."""
    rng = np.random.default_rng(seed)
    sizes = [(5, 64), (64, 64), (64, 64), (64, 64), (64, 2)]
    out = []
    for li, (fin, fout) in enumerate(sizes):
        lim = np.sqrt(6.0 / fin) if li < 4 else np.sqrt(1.0 / fin) * 0.1
        out.append(rng.uniform(-lim, lim, (fout, fin)).ravel())
        out.append(np.zeros(fout))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == _abi.MLP_PARAMS
    return p


def set_precision(bf16: bool) -> None:
    """f32 (False) or the fused inference's bf16 numerics (True) for the
    library's forward / gradient / training calls."""
    lib = _abi.load_library()
    _check(lib, lib.pstm_mlp_set_precision(1 if bf16 else 0), "pstm_mlp_set_precision")


def vti_leg_times(tau, h, v, eta):
    """Exact one-way VTI times on the device (f64 torch tensors)."""
    import torch

    lib = _abi.load_library()
    tau, h, v, eta = (a.to(torch.float64).contiguous() for a in (tau, h, v, eta))
    out = torch.empty_like(tau)
    _check(lib, lib.pstm_vti_leg_times(_p(tau), _p(h), _p(v), _p(eta), int(tau.numel()), _p(out),
                                       _stream()), "pstm_vti_leg_times")
    return out


def hyperbolic(tau, d, v):
    return (tau * tau + (d / v) ** 2).sqrt()


@dataclass
class TrainingSet:
    x_raw: object      # (n, 5) f64: tau, d_s, d_r, v, eta
    x: object          # (n, 5) f32 normalised
    y: object          # (n, 2) f32: out_scale * (c_s, c_r)
    t_exact: object    # (n, 2) f64 exact leg times


def make_training_set(vcube, ecube, n: int, d_max: float, norm: Normalizer, seed: int,
                      device="cuda", t_max_ms: float | None = None) -> TrainingSet:
    """n samples on the job's own model: x, y uniform over the cube's extent,
    t0 uniform over [0, t_max], d_s, d_r uniform over [0, d_max]; v and eta by
    trilinear interpolation of the cubes.  Targets out_scale * g per leg, g =
    c / eta (anisotropic_factor per unit eta); samples with eta < 1e-6 are
    left out (fewer than n where the model reaches eta = 0)."""
    import torch

    rng = np.random.default_rng(seed)
    xs = rng.uniform(vcube.x_axis[0], vcube.x_axis[-1], n)
    ys = rng.uniform(vcube.y_axis[0], vcube.y_axis[-1], n)
    tmax = t_max_ms if t_max_ms is not None else vcube.t_axis_ms[-1]
    ts = rng.uniform(0.0, tmax, n)
    ds = rng.uniform(0.0, d_max, n)
    dr = rng.uniform(0.0, d_max, n)
    v = torch.as_tensor(_trilinear(vcube, xs, ys, ts), device=device)
    e = torch.as_tensor(_trilinear(ecube, xs, ys, ts), device=device)
    tau = torch.as_tensor(ts / 2000.0, device=device)
    dst = torch.as_tensor(ds, device=device)
    drt = torch.as_tensor(dr, device=device)
    t_s = vti_leg_times(tau, dst, v, e)
    t_r = vti_leg_times(tau, drt, v, e)
    # targets g = c / eta; where eta < 1e-6 the time does not depend on g
    # (and c / eta is not defined at eta = 0): those samples are left out
    keep = e >= 1e-6
    tau, dst, drt, v, e, t_s, t_r = (a[keep] for a in (tau, dst, drt, v, e, t_s, t_r))
    xr = torch.stack([tau, dst, drt, v, e], 1)
    g = torch.stack([anisotropic_factor(t_s, tau, dst, v, e), anisotropic_factor(t_r, tau, drt, v, e)], 1) / e[:, None]
    y = (g * norm.out_scale).float().contiguous()
    return TrainingSet(xr, norm.apply(xr), y, torch.stack([t_s, t_r], 1))


def anisotropic_factor(t, tau, d, v, eta):
    """c = (t^2 - tau^2) / (d / v)^2 - 1 from exact leg times t (f64).  At
    tau = 0 the exact time is 0 by convention, and c takes its limit there,
    -2 eta / (1 + 2 eta); where (d / v)^2 is below 1e-10 s^2 the time does not
    depend on c and c = 0."""
    import torch

    hs = (d / v) ** 2
    c = (t * t - tau * tau) / hs.clamp_min(1e-300) - 1.0
    c = torch.where(tau > 0, c, -2.0 * eta / (1.0 + 2.0 * eta))
    return torch.where(hs > 1e-10, c, torch.zeros_like(c))


def learned_leg_time(tau, d, v, c):
    """sqrt(tau^2 + (d / v)^2 max(1 + c, 0)), as the gather evaluates it."""
    return (tau * tau + (d / v) ** 2 * (1.0 + c).clamp_min(0.0)).sqrt()


def curved_velocity_at(tau, v0: float, k: float):
    """v(z) at the depth whose vertical one-way time is tau in
    V(z) = v0 + k z: z = (v0 / k)(e^{k tau} - 1) (curved_ray.py:239-262),
    v(z) = v0 + k z (f64 tensors)."""
    import torch

    tau = torch.as_tensor(tau, dtype=torch.float64)
    if abs(k) < 1e-10:
        return torch.full_like(tau, float(v0))
    z = (v0 / k) * (torch.exp(k * tau) - 1.0)
    return v0 + k * z


def curved_leg_times(tau, h, v0: float, k: float):
    """One-way curved-ray time of the reference (curved_ray.py:19-52, the
    depth from tau by :239-262) as f64 tensor arithmetic: the exact targets
    of the learned curved traveltime."""
    import torch

    tau = torch.as_tensor(tau, dtype=torch.float64)
    h = torch.as_tensor(h, dtype=torch.float64)
    if abs(k) < 1e-10:
        z = v0 * tau
        return (h * h + z * z).sqrt() / v0
    vz = curved_velocity_at(tau, v0, k)
    kx = k * h
    t1 = vz + (vz * vz + kx * kx).sqrt()
    t2 = v0 + (v0 * v0 + kx * kx).sqrt()
    return torch.log(t1 / t2) / k


def curved_depth_and_vg(tau, v0: float, k: float):
    """Depth z of the vertical one-way time tau in V(z) = v0 + k z and the
    geometric-mean velocity vg = sqrt(v0 v(z)) of the learned curved leg."""
    import torch

    vz = curved_velocity_at(tau, v0, k)
    return (vz - v0) / k, (v0 * vz).sqrt() if isinstance(vz, torch.Tensor) else np.sqrt(v0 * vz)


def learned_leg_time_curved(tau, d, c, v0: float, k: float):
    """z / sqrt(vg^2 + k^2 d^2) max(1 + c, 0), as the gather evaluates the
    learned curved leg.  The reference's leg (1/k) ln[(v(z) + sqrt(v(z)^2 +
    k^2 d^2)) / (v0 + sqrt(v0^2 + k^2 d^2))] equals z / sqrt(xi^2 + k^2 d^2)
    for some xi between v0 and v(z) (mean value theorem on asinh), so c is a
    small correction of the geometric-mean choice xi = vg."""
    z, vg = curved_depth_and_vg(tau, v0, k)
    return z / (vg * vg + (k * d) ** 2).sqrt() * (1.0 + c).clamp_min(0.0)


def make_training_set_curved(v0: float, k: float, n: int, d_max: float, norm: Normalizer, seed: int,
                             t_max_ms: float, device="cuda") -> TrainingSet:
    """n samples of the job's V(z) = v0 + k z: t0 uniform over [0, t_max],
    d_s, d_r uniform over [0, d_max]; raw inputs (tau, d_s, d_r, vg, k),
    encoded with the curved sines k d / sqrt(vg^2 + k^2 d^2); targets
    out_scale * c with c = t / (z / sqrt(vg^2 + k^2 d^2)) - 1 per leg (0 at
    t0 = 0, where both times are 0)."""
    import torch

    rng = np.random.default_rng(seed)
    tau = torch.as_tensor(rng.uniform(0.0, t_max_ms, n) / 2000.0, device=device)
    dst = torch.as_tensor(rng.uniform(0.0, d_max, n), device=device)
    drt = torch.as_tensor(rng.uniform(0.0, d_max, n), device=device)
    z, vg = curved_depth_and_vg(tau, v0, k)
    kk = torch.full_like(tau, float(k))
    t_s = curved_leg_times(tau, dst, v0, k)
    t_r = curved_leg_times(tau, drt, v0, k)

    def factor(t, d):
        tref = z / (vg * vg + (k * d) ** 2).sqrt()
        return torch.where(tref > 0, t / tref.clamp_min(1e-300) - 1.0, torch.zeros_like(tref))

    xr = torch.stack([tau, dst, drt, vg, kk], 1)
    c = torch.stack([factor(t_s, dst), factor(t_r, drt)], 1)
    return TrainingSet(xr, norm.encode_curved(tau, dst, drt, vg, kk), (c * norm.out_scale).float().contiguous(),
                       torch.stack([t_s, t_r], 1))


def _trilinear(cube, x, y, t):
    """Trilinear interpolation of a VelocityCube-like grid at scattered points
    (numpy, clamped), vectorised."""
    def idx(axis, p):
        axis = np.asarray(axis, dtype=np.float64)
        p = np.clip(p, axis[0], axis[-1])
        i = np.clip(np.searchsorted(axis, p, side="right") - 1, 0, len(axis) - 2)
        return i, (p - axis[i]) / (axis[i + 1] - axis[i])

    ix, fx = idx(cube.x_axis, x)
    iy, fy = idx(cube.y_axis, y)
    it, ft = idx(cube.t_axis_ms, t)
    d = cube.data
    out = np.zeros(len(x))
    for a in (0, 1):
        wa = fx if a else 1 - fx
        for b in (0, 1):
            wb = fy if b else 1 - fy
            for c in (0, 1):
                wc = ft if c else 1 - ft
                out += wa * wb * wc * d[ix + a, iy + b, it + c]
    return out


@dataclass
class TraveltimeSurrogate:
    """Parameters (device f32), Adam moments and the input normaliser."""

    norm: Normalizer
    params: object = None
    m: object = None
    v: object = None
    grads: object = None
    step: int = 0
    history: list = field(default_factory=list)
    # numerics of the fused inference: 1 = weights rounded to bf16 (train with
    # a quantisation-aware tail), 2 = each weight as two bf16 terms (an
    # f32-trained network keeps about 16 bits)
    weight_terms: int = 1
    # traveltime the network learns: "vti" (inputs tau, s_s, s_r, v_rms,
    # eta) or "curved" (tau, s_s, s_r, vg, k; PSTM_TT_CURVED_LEARNED) with
    # the job's surface velocity curved_v0
    kind: str = "vti"
    curved_v0: float = 0.0

    @classmethod
    def create(cls, norm: Normalizer, seed: int = 0, device="cuda") -> "TraveltimeSurrogate":
        import torch

        p = torch.as_tensor(init_params(seed), device=device)
        return cls(norm, p, torch.zeros_like(p), torch.zeros_like(p), torch.zeros_like(p))

    def forward(self, x, bf16: bool = False):
        """Network output (n, 2) for normalised inputs x (n, 5) f32, in f32 or
        with the fused inference's bf16 numerics."""
        import torch

        lib = _abi.load_library()
        x = x.float().contiguous()
        y = torch.empty((x.shape[0], 2), dtype=torch.float32, device=x.device)
        set_precision(bf16)
        try:
            _check(lib, lib.pstm_mlp_forward(_p(self.params), _p(x), int(x.shape[0]), _p(y), _stream()),
                   "pstm_mlp_forward")
        finally:
            set_precision(False)
        return y

    def forward_fused(self, x):
        """Network output (n, 2) through the migration kernel's own fused
        inference code (BF16 MFMA, f32 accumulation, this surrogate's
        weight_terms), for normalised inputs x (n, 5)."""
        import torch

        lib = _abi.load_library()
        x = x.float().contiguous()
        y = torch.empty((x.shape[0], 2), dtype=torch.float32, device=x.device)
        rc = lib.pstm_mlp_forward_mfma_terms(_p(self.params), _p(x), int(x.shape[0]), _p(y), int(self.weight_terms),
                                             _stream())
        if rc != 0:
            raise (ValueError if rc == 1 else RuntimeError)(f"pstm_mlp_forward_mfma_terms: {_abi.last_error(lib)}")
        return y

    def loss_grad(self, x, y):
        import torch

        lib = _abi.load_library()
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        _check(lib, lib.pstm_mlp_loss_grad(_p(self.params), _p(x), _p(y), int(x.shape[0]), _p(self.grads),
                                           _p(loss), _stream()), "pstm_mlp_loss_grad")
        return loss, self.grads

    def train(self, ts: TrainingSet, steps: int, batch: int = 8192, lr: float = 2e-3,
              lr_final: float = 5e-5, seed: int = 0, log_every: int = 0, qat_fraction: float = 0.0,
              verbose: bool = False):
        """Adam with a cosine-decayed learning rate over `steps` minibatches;
        the last qat_fraction of the steps train through the bf16 numerics of
        the fused inference (quantisation-aware).  The whole run is one
        library call (pstm_mlp_train): one fused forward/backward kernel and
        one Adam kernel per step, minibatches drawn on the device, replayed
        from a captured hipGraph.  Every log_every steps the loss is recorded
        in history (and printed to stderr when verbose)."""
        import sys

        import torch

        lib = _abi.load_library()
        q_from = steps - int(round(qat_fraction * steps))
        n = int(ts.x.shape[0])
        x = ts.x.float().contiguous()
        y = ts.y.float().contiguous()
        le = int(log_every) if log_every else 0
        slots = (steps + le - 1) // le + 1 if le else 0
        log = torch.zeros(max(slots, 1), dtype=torch.float32, device=x.device)
        _check(lib, lib.pstm_mlp_train(_p(self.params), _p(self.m), _p(self.v), _p(x), _p(y), n, int(steps),
                                       int(batch), float(lr), float(lr_final), 0.9, 0.999, 1e-8, int(self.step),
                                       int(q_from), int(seed) & (2**64 - 1), le, _p(log) if le else None,
                                       _stream()), "pstm_mlp_train")
        self.step += steps
        if le:
            vals = log.cpu().numpy()
            its = list(range(0, steps, le))
            if its[-1] != steps - 1:
                its.append(steps - 1)
            for j, i in enumerate(its):
                self.history.append((i, float(vals[j])))
                if verbose:
                    print(f"[surrogate] step {i}: loss {vals[j]:.3e}", file=sys.stderr, flush=True)
        return self

    def train_stepwise(self, ts: TrainingSet, steps: int, batch: int = 8192, lr: float = 2e-3,
                       lr_final: float = 5e-5, seed: int = 0, log_every: int = 0, qat_fraction: float = 0.0,
                       verbose: bool = False):
        """The same schedule as ``train`` with one host-driven library call
        per step (pstm_mlp_train_step: per-layer kernels, minibatches drawn
        by torch): the reference for the fused training."""
        import sys

        import torch

        lib = _abi.load_library()
        q_from = steps - int(round(qat_fraction * steps))
        n = int(ts.x.shape[0])
        g = torch.Generator(device=ts.x.device)
        g.manual_seed(seed)
        loss = torch.zeros(1, dtype=torch.float32, device=ts.x.device)
        for i in range(steps):
            idx = torch.randint(0, n, (batch,), device=ts.x.device, generator=g)
            xb = ts.x.index_select(0, idx).contiguous()
            yb = ts.y.index_select(0, idx).contiguous()
            self.step += 1
            a = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * i / max(steps - 1, 1)))
            set_precision(i >= q_from)
            try:
                _check(lib, lib.pstm_mlp_train_step(_p(self.params), _p(self.m), _p(self.v), _p(self.grads),
                                                    _p(xb), _p(yb), batch, float(a), 0.9, 0.999, 1e-8,
                                                    self.step, _p(loss), _stream()), "pstm_mlp_train_step")
            finally:
                set_precision(False)
            if log_every and (i % log_every == 0 or i == steps - 1):
                self.history.append((i, float(loss.item())))
                if verbose:
                    print(f"[surrogate] step {i}: loss {self.history[-1][1]:.3e}", file=sys.stderr, flush=True)
        return self

    def leg_times(self, ts: TrainingSet, bf16: bool = False):
        """Predicted (t_s, t_r) f64 for a sample set: f32 numerics, or (bf16)
        those of the fused inference."""
        import torch

        y = self.forward_fused(ts.x) if bf16 else self.forward(ts.x)
        c = y.double() / self.norm.out_scale
        tau, ds, dr, v = ts.x_raw[:, 0], ts.x_raw[:, 1], ts.x_raw[:, 2], ts.x_raw[:, 3]
        if self.kind == "curved":
            # raw inputs (tau, d_s, d_r, vg, k): z = (vg^2 / v0 - v0) / k
            kg = ts.x_raw[:, 4]
            z = (v * v / self.curved_v0 - self.curved_v0) / kg

            def leg(d, cc):
                return z / (v * v + (kg * d) ** 2).sqrt() * (1.0 + cc).clamp_min(0.0)

            return torch.stack([leg(ds, c[:, 0]), leg(dr, c[:, 1])], 1)
        # VTI: the outputs are g = c / eta
        eta = ts.x_raw[:, 4]
        return torch.stack([learned_leg_time(tau, ds, v, eta * c[:, 0]), learned_leg_time(tau, dr, v, eta * c[:, 1])],
                           1)

    def error_samples(self, ts: TrainingSet, dt_ms: float, bf16: bool = False) -> dict:
        """Error of t_s + t_r against the exact sum, in input samples."""
        pred = self.leg_times(ts, bf16).sum(1)
        exact = ts.t_exact.sum(1)
        err = (pred - exact).abs() / (dt_ms / 1000.0)
        return {"max": float(err.max()), "p99": float(err.quantile(0.99)) if err.numel() <= 16_000_000 else None,
                "mean": float(err.mean())}

    def bf16_params(self):
        """Parameters rounded to bfloat16 (the precision the migration uses)."""
        import torch

        return self.params.to(torch.bfloat16)


def node_derivative(f, ok, x, i):
    """Derivative at node i (1 or 2) of values f[0..3] at positions x[0..3]
    (only the nodes with ok set), as the gather forms it: the three-point
    difference over i - 1, i, i + 1 where both neighbours are there, else the
    one-sided three-point difference over i, i -+ 1, i -+ 2, else the
    two-point one; all second-order, exact for a quadratic (f64 tensors)."""
    import torch

    def d3(a, b, c, xa, xb, xc):
        # derivative at xa of the parabola through (xa, a), (xb, b), (xc, c)
        h1, h2 = xb - xa, xc - xb
        return -(2.0 * h1 + h2) / (h1 * (h1 + h2)) * a + (h1 + h2) / (h1 * h2) * b - h1 / (h2 * (h1 + h2)) * c

    lo, hi = i - 1, i + 1
    h1, h2 = x[i] - x[lo], x[hi] - x[i]
    central = (h1 * h1 * (f[hi] - f[i]) + h2 * h2 * (f[i] - f[lo])) / (h1 * h2 * (h1 + h2))
    if i == 1:
        one = torch.where(ok[3], d3(f[1], f[2], f[3], x[1], x[2], x[3]), (f[2] - f[1]) / (x[2] - x[1]))
        return torch.where(ok[0], central, one)
    # backward: the parabola through nodes 2, 1, 0 (negative spacings)
    one = torch.where(ok[0], d3(f[2], f[1], f[0], x[2], x[1], x[0]), (f[2] - f[1]) / (x[2] - x[1]))
    return torch.where(ok[3], central, one)


# pillars per workgroup of the fused gather: the network is evaluated at
# three pillars of each group of consecutive (tile-local, row-major) pillars
# and interpolated quadratically in between
GROUP_PILLARS = 8


def pillar_group_ends(pil, npil: int, group: int = GROUP_PILLARS):
    """For tile-local pillar indices pil (int64 tensor) of a tile of npil
    pillars: the pillars of each one's group at which the gather evaluates the
    network, (n, 3) = (first, middle, last valid pillar of the group), and the
    weights (n, 3) that make the pillar's value from theirs, as the gather
    interpolates the network's outputs: the quadratic through the three in
    the pillar index (a leg's squared distance is quadratic in it), the line
    through two for a group of two pillars, the value itself for one."""
    import torch

    e0 = (pil // group) * group
    p_hi = torch.clamp(npil - 1 - e0, max=group - 1)
    p_mid = (p_hi + 1) // 2
    p = (pil - e0).double()
    a, b, c = torch.zeros_like(p), p_mid.double(), p_hi.double()
    quad = p_hi >= 2
    bs, cs = torch.where(quad, b, 1.0), torch.where(quad, c, 2.0)
    w0 = (p - bs) * (p - cs) / ((a - bs) * (a - cs))
    w1 = (p - a) * (p - cs) / ((bs - a) * (bs - cs))
    w2 = (p - a) * (p - bs) / ((cs - a) * (cs - bs))
    lin = p / c.clamp_min(1.0)
    W = torch.stack([torch.where(quad, w0, torch.where(p_hi == 1, 1.0 - lin, 1.0)),
                     torch.where(quad, w1, 0.0),
                     torch.where(quad, w2, torch.where(p_hi == 1, lin, 0.0))], 1)
    E = torch.stack([e0, e0 + p_mid, e0 + p_hi], 1)
    return E, W


def lattice_times(sur: TraveltimeSurrogate, tau_axis, k, ds, dr, v_at, eta_at, node_spacing, ends=None):
    """Learned two-leg time of output samples k as the gather forms it.  The
    network (the fused inference's own code) gives the legs' factors g at the
    lattice nodes (pm_amd.lattice; ``node_spacing`` a Lattice or an int, the
    uniform spacing) for three pillars of the sample pillar's group,
    interpolated quadratically in the pillar index (``ends``); along t0 the
    sample's segment n_j <= k < n_{j+1} takes the cubic Hermite polynomial in
    the fraction (k - n_j) / (n_{j+1} - n_j) through g at n_j and n_{j+1}
    with slopes by node_derivative over nodes j - 1 .. j + 2.  The legs then
    use the sample's own v and eta: VTI sqrt(tau^2 + (d / v)^2 max(1 + eta g,
    0)) (t = t0 where t0 <= 0); curved ray z / sqrt(vg^2 + k^2 d^2) max(1 + c,
    0).

    tau_axis: (nt,) one-way times t0 / 2 (f64); k: (n,) int64 sample indices;
    ds, dr: (n,) leg distances; v_at(kk), eta_at(kk): (n,) f64 values at the
    sample's own pillar and samples kk.  ends: None (the sample's own
    pillar) or a dict with 'ds', 'dr' ((n, 3) leg distances from the group's
    three evaluated pillars), 'v_at', 'eta_at' (callables (kk, e) -> (n,)
    values at pillar e) and 'w' ((n, 3) weights, pillar_group_ends)."""
    import torch

    from pm_amd.lattice import as_lattice

    nt = int(tau_axis.shape[0])
    nodes = torch.as_tensor(as_lattice(node_spacing).nodes(nt), device=k.device)
    nj = int(nodes.shape[0])
    j = (torch.searchsorted(nodes, k, right=True) - 1).clamp(0, nj - 2)
    if ends is None:
        ends = {"ds": ds[:, None], "dr": dr[:, None], "v_at": lambda kk, e: v_at(kk),
                "eta_at": lambda kk, e: eta_at(kk), "w": torch.ones_like(ds)[:, None]}

    def g_at(kn, e):
        """Network outputs (n, 2) at nodes kn for end pillar e."""
        tau = tau_axis[kn]
        v, et = ends["v_at"](kn, e), ends["eta_at"](kn, e)
        enc = sur.norm.encode_curved if sur.kind == "curved" else sur.norm.encode
        x = enc(tau, ends["ds"][:, e], ends["dr"][:, e], v, et)
        return sur.forward_fused(x).double() / sur.norm.out_scale

    # the four nodes j - 1 .. j + 2 (positions of off-axis ones 1 apart)
    jj = [j - 1, j, j + 1, j + 2]
    ok = [(q >= 0) & (q < nj) for q in jj]
    x = [nodes[q.clamp(0, nj - 1)].double() + torch.where(q < 0, q.double(), 0.0) +
         torch.where(q >= nj, (q - nj + 1).double(), 0.0) for q in jj]
    W = ends["w"]
    g = []
    for q in jj:
        kn = nodes[q.clamp(0, nj - 1)]
        g.append(sum(W[:, e:e + 1] * g_at(kn, e) for e in range(W.shape[1])))
    L = x[2] - x[1]
    f = (k - nodes[j]).double() / L
    gk = []
    for leg in (0, 1):
        gl = [gn[:, leg] for gn in g]
        b0 = L * node_derivative(gl, ok, x, 1)
        b1 = L * node_derivative(gl, ok, x, 2)
        gk.append(gl[1] + f * (b0 + f * ((3.0 * (gl[2] - gl[1]) - 2.0 * b0 - b1) + f * (2.0 * (gl[1] - gl[2]) + b0 + b1))))
    tau = tau_axis[k]
    v = v_at(k)
    if sur.kind == "curved":
        # v_at is vg at the sample, eta_at the gradient k; t = 0 at t0 = 0
        kg = float(eta_at(k[:1])[0])
        vg = v
        z = (vg * vg / float(sur.curved_v0) - float(sur.curved_v0)) / kg
        leg = lambda d, cc: z / (vg * vg + (kg * d) ** 2).sqrt() * (1.0 + cc).clamp_min(0.0)  # noqa: E731
        return leg(ds, gk[0]) + leg(dr, gk[1])
    eta = eta_at(k)
    t = learned_leg_time(tau, ds, v, eta * gk[0]) + learned_leg_time(tau, dr, v, eta * gk[1])
    return torch.where(tau > 0, t, 2.0 * tau)
