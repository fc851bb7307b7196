"""Traveltime surrogate: the FP64 NumPy network (oracle/mlp.py) pinned by
finite differences and the closed-form Adam update, its bf16 rounding pinned
against torch's; the HIP forward, gradient and Adam kernels and the fused
path's BF16 MFMA inference checked against it on the GPU; exact VTI targets
against the oracle's golden-pinned VTI time; training reaches the
0.25-sample bar under the fused inference's numerics."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as omlp
from pm_amd.kernels import _abi


def _rand_params(seed=0):
    from pm_amd.surrogate import init_params

    p = init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    return p + 0.01 * rng.standard_normal(p.size)  # non-zero biases too


def test_layout_matches_header():
    assert sum(fi * fo + fo for fi, fo in omlp.SIZES) == _abi.MLP_PARAMS == 12994


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(11)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, -3.0]]).astype(np.float32)
    ref = torch.as_tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(omlp.bf16_round(x), ref)


def test_bf16_pair_carries_sixteen_bits():
    rng = np.random.default_rng(12)
    x = rng.uniform(-8, 8, 10000).astype(np.float32)
    err = np.abs(omlp.bf16_pair(x) - x.astype(np.float64))
    assert np.all(err <= 2.0 ** -16 * np.abs(x.astype(np.float64)))


def test_bf16_forward_close_to_f64():
    """The bf16 numerics stay near the f64 network (weights rounded: a few
    1e-3 relative), and equal it for weights that are bf16 already."""
    rng = np.random.default_rng(13)
    p = _rand_params(13)
    x = rng.uniform(-1, 1, (256, 5))
    y64, yb = omlp.forward(p, x), omlp.forward_bf16(p, x)
    assert np.max(np.abs(yb - y64)) < 0.05 * np.max(np.abs(y64))
    pb = omlp.bf16_round(p.astype(np.float32))
    xb = omlp.bf16_round(x.astype(np.float32))
    np.testing.assert_allclose(omlp.forward_bf16(pb, xb), omlp.forward(pb, xb), rtol=2e-4, atol=1e-5)


def test_scaled_bf16_forward_is_the_same_network():
    """The fused inference's scaled evaluation (every value times
    c = -log2(e), SiLU as z' / (1 + 2^z'), outputs divided by c) is the same
    network: with bf16 weights it equals the f64 forward as closely as the
    unscaled bf16 numerics do, and it differs from the unscaled numerics only
    by the rounding of the activation pairs (a dropped 1 / c, a sign or the
    bias scale would be off by O(1))."""
    rng = np.random.default_rng(15)
    p = _rand_params(15)
    x = rng.uniform(-1, 1, (256, 5))
    pb = omlp.bf16_round(p.astype(np.float32))
    y64 = omlp.forward(pb, x)
    ys = omlp.forward_bf16(pb, x, omlp.SILU_C)
    scale = np.max(np.abs(y64))
    assert np.max(np.abs(ys - y64)) < 2e-4 * scale
    assert np.max(np.abs(ys - omlp.forward_bf16(pb, x))) < 1e-4 * scale
    # split weights, scaled and unscaled
    ys2 = omlp.forward_bf16_split(p, x)
    assert np.max(np.abs(ys2 - omlp.forward_bf16_split(p, x, 1.0))) < 1e-4 * np.max(np.abs(ys2))
    assert np.max(np.abs(ys2 - omlp.forward(p, x))) < 2e-4 * np.max(np.abs(ys2))


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _rand_params(3)
    x = rng.uniform(-1, 1, (7, 5))
    t = rng.uniform(-0.1, 0.1, (7, 2))
    loss, g = omlp.loss_grad(p, x, t)
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (omlp.loss_grad(pp, x, t)[0] - omlp.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, i
    assert loss == pytest.approx(np.mean((omlp.forward(p, x) - t) ** 2))


def test_adam_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.1, -0.3, 0.0])
    p1, m, v = omlp.adam(p, g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-8, 1)
    # first step with bias correction: update = lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


@pytest.mark.gpu
class TestHipSurrogate:
    def test_forward_and_gradient_match_fp64(self):
        import torch

        from pm_amd.surrogate import TraveltimeSurrogate, Normalizer

        rng = np.random.default_rng(5)
        p = _rand_params(5)
        x = rng.uniform(-1, 1, (1000, 5))
        t = rng.uniform(-0.1, 0.05, (1000, 2))
        s = TraveltimeSurrogate.create(Normalizer(2.0, 3000.0, 1500.0, 4500.0, 0.2))
        s.params = torch.as_tensor(p.astype(np.float32), device="cuda")
        xd = torch.as_tensor(x, dtype=torch.float32, device="cuda")
        td = torch.as_tensor(t, dtype=torch.float32, device="cuda")
        y = s.forward(xd).cpu().numpy()
        yr = omlp.forward(p.astype(np.float32).astype(np.float64), x.astype(np.float32))
        err_f = np.abs(y - yr).max() / np.abs(yr).max()
        loss, g = s.loss_grad(xd, td)
        lr_, gr = omlp.loss_grad(p.astype(np.float32).astype(np.float64), x.astype(np.float32),
                                 t.astype(np.float32))
        g = g.cpu().numpy()
        err_g = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        print(f"forward rel {err_f:.2e}, gradient rel-L2 {err_g:.2e}, loss {float(loss.item()):.6e} vs {lr_:.6e}")
        assert err_f < 1e-5 and err_g < 1e-4
        assert abs(float(loss.item()) - lr_) <= 1e-5 * lr_

    def test_adam_matches(self):
        import ctypes

        import torch

        lib = _abi.load_library()
        rng = np.random.default_rng(6)
        p = rng.standard_normal(_abi.MLP_PARAMS)
        g = rng.standard_normal(_abi.MLP_PARAMS)
        m = 0.1 * rng.standard_normal(_abi.MLP_PARAMS)
        v = np.abs(rng.standard_normal(_abi.MLP_PARAMS))
        dev = [torch.as_tensor(a.astype(np.float32), device="cuda") for a in (p, g, m, v)]
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib.pstm_mlp_adam(*(d.data_ptr() for d in dev), 1e-3, 0.9, 0.999, 1e-8, 7, s) == 0
        pr, mr, vr = omlp.adam(p.astype(np.float32).astype(np.float64), g.astype(np.float32),
                               m.astype(np.float32), v.astype(np.float32), 1e-3, 0.9, 0.999, 1e-8, 7)
        np.testing.assert_allclose(dev[0].cpu().numpy(), pr, rtol=0, atol=1e-6)
        np.testing.assert_allclose(dev[2].cpu().numpy(), mr, rtol=1e-6, atol=1e-7)

    def test_exact_targets_match_oracle(self):
        import torch

        from oracle import oracle as orc
        from pm_amd.surrogate import vti_leg_times

        rng = np.random.default_rng(8)
        tau, h = rng.uniform(0, 2, 500), rng.uniform(0, 4000, 500)
        v, e = rng.uniform(1500, 4500, 500), rng.uniform(0, 0.2, 500)
        tau[:5] = 0.0
        got = vti_leg_times(*(torch.as_tensor(a, device="cuda") for a in (tau, h, v, e))).cpu().numpy()
        ref = np.array([orc.vti_leg(*a) for a in zip(tau, h, v, e)])
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)

    def test_bf16_mode_and_mfma_match_oracle(self):
        """pstm_mlp_forward in bf16 mode and the fused path's MFMA inference
        (pstm_mlp_forward_mfma) against the bf16 NumPy forward; n = 1000 is
        not a multiple of the 16-row MFMA tile."""
        import torch

        from pm_amd.surrogate import _stream

        lib = _abi.load_library()
        rng = np.random.default_rng(14)
        p = _rand_params(14).astype(np.float32)
        x = rng.uniform(-1, 1, (1000, 5)).astype(np.float32)
        ref = omlp.forward_bf16(p, x)  # the trainer's bf16 mode
        ref_c = omlp.forward_bf16(p, x, omlp.SILU_C)  # the fused inference (scaled)
        pd, xd = torch.as_tensor(p, device="cuda"), torch.as_tensor(x, device="cuda")
        scale = np.max(np.abs(ref))
        ref_split = omlp.forward_bf16_split(p, x)
        for mode in ("bf16", "mfma", "mfma_split"):
            y = torch.full((1000, 2), float("nan"), device="cuda")
            if mode == "mfma_split":
                assert lib.pstm_mlp_forward_mfma_terms(pd.data_ptr(), xd.data_ptr(), 1000, y.data_ptr(), 2,
                                                       _stream()) == 0
                err = np.max(np.abs(y.cpu().numpy() - ref_split)) / scale
                print(mode, "max abs error / max |y|", err, "(bf16 weights:", np.max(np.abs(ref - ref_split)) / scale,
                      ")")
                assert err < 2e-5
                continue
            if mode == "bf16":
                assert lib.pstm_mlp_set_precision(1) == 0
                try:
                    assert lib.pstm_mlp_forward(pd.data_ptr(), xd.data_ptr(), 1000, y.data_ptr(), _stream()) == 0
                finally:
                    lib.pstm_mlp_set_precision(0)
            else:
                assert lib.pstm_mlp_forward_mfma(pd.data_ptr(), xd.data_ptr(), 1000, y.data_ptr(), _stream()) == 0
            got = y.cpu().numpy()
            err = np.max(np.abs(got - (ref if mode == "bf16" else ref_c))) / scale
            print(mode, "max abs error / max |y|", err)
            assert err < 2e-5, mode

    def test_fused_gradient_matches_fp64(self):
        """The fused training kernel's loss and gradient (forward, loss and
        backward of 32-sample chunks in LDS, per-workgroup partials) against
        the FP64 NumPy network; n = 9,999 is not a multiple of 32 and spans
        more chunks than workgroups."""
        import torch

        from pm_amd.surrogate import _stream

        lib = _abi.load_library()
        rng = np.random.default_rng(15)
        p = _rand_params(15)
        n = 9999
        x = rng.uniform(-1, 1, (n, 5))
        t = rng.uniform(-0.1, 0.05, (n, 2))
        pd = torch.as_tensor(p.astype(np.float32), device="cuda")
        xd = torch.as_tensor(x, dtype=torch.float32, device="cuda")
        td = torch.as_tensor(t, dtype=torch.float32, device="cuda")
        g = torch.full((_abi.MLP_PARAMS,), float("nan"), device="cuda")
        loss = torch.zeros(1, device="cuda")
        assert lib.pstm_mlp_loss_grad_fused(pd.data_ptr(), xd.data_ptr(), td.data_ptr(), n, g.data_ptr(),
                                            loss.data_ptr(), _stream()) == 0
        lr_, gr = omlp.loss_grad(p.astype(np.float32).astype(np.float64), x.astype(np.float32), t.astype(np.float32))
        g = g.cpu().numpy()
        err_g = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        # per layer too: a dropped term in one small layer would hide in the total
        offs = np.cumsum([0] + [a * b + b for a, b in [(5, 64), (64, 64), (64, 64), (64, 64), (64, 2)]])
        per = [np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b]) for a, b in zip(offs[:-1], offs[1:])]
        print(f"fused gradient rel-L2 {err_g:.2e}, per layer {[f'{e:.1e}' for e in per]}, loss "
              f"{float(loss.item()):.6e} vs {lr_:.6e}")
        assert err_g < 1e-4 and max(per) < 1e-4
        assert abs(float(loss.item()) - lr_) <= 1e-5 * lr_

    def test_fused_gradient_bf16_mode_matches_stepwise(self):
        """Quantisation-aware numerics: the fused gradient against the
        per-layer kernels' gradient in the same bf16 mode (straight-through;
        the fused kernel's layer-0 weight gradient uses the input as the bf16
        forward saw it, so the two agree to bf16 input rounding there)."""
        import torch

        from pm_amd.surrogate import _stream

        lib = _abi.load_library()
        rng = np.random.default_rng(16)
        p = _rand_params(16).astype(np.float32)
        n = 4096
        pd = torch.as_tensor(p, device="cuda")
        xd = torch.as_tensor(rng.uniform(-1, 1, (n, 5)).astype(np.float32), device="cuda")
        td = torch.as_tensor(rng.uniform(-0.1, 0.05, (n, 2)).astype(np.float32), device="cuda")
        outs = []
        assert lib.pstm_mlp_set_precision(1) == 0
        try:
            for fn in (lib.pstm_mlp_loss_grad_fused, lib.pstm_mlp_loss_grad):
                g = torch.zeros(_abi.MLP_PARAMS, device="cuda")
                loss = torch.zeros(1, device="cuda")
                assert fn(pd.data_ptr(), xd.data_ptr(), td.data_ptr(), n, g.data_ptr(), loss.data_ptr(),
                          _stream()) == 0
                outs.append((g.cpu().numpy(), float(loss.item())))
        finally:
            lib.pstm_mlp_set_precision(0)
        (gf, lf), (gs, ls) = outs
        rel = np.linalg.norm(gf - gs) / np.linalg.norm(gs)
        print(f"bf16-mode fused vs per-layer gradient rel-L2 {rel:.2e}, loss {lf:.6e} vs {ls:.6e}")
        assert rel < 1e-2 and abs(lf - ls) <= 1e-5 * ls

    def test_fused_training_step_is_adam_on_the_gradient(self):
        """One iteration of pstm_mlp_train on a one-row set (every drawn
        sample is row 0, so the minibatch gradient is that row's): the
        parameters move by the Adam closed form with the schedule's first
        learning rate."""
        import torch

        from pm_amd.surrogate import _stream

        lib = _abi.load_library()
        rng = np.random.default_rng(17)
        p = _rand_params(17).astype(np.float32)
        x = rng.uniform(-1, 1, (1, 5))
        t = rng.uniform(-0.1, 0.05, (1, 2))
        _, gr = omlp.loss_grad(p.astype(np.float64), x.astype(np.float32), t.astype(np.float32))
        pd = torch.as_tensor(p, device="cuda")
        m = torch.zeros_like(pd)
        v = torch.zeros_like(pd)
        xd = torch.as_tensor(x.astype(np.float32), device="cuda")
        td = torch.as_tensor(t.astype(np.float32), device="cuda")
        log = torch.full((2,), float("nan"), device="cuda")
        assert lib.pstm_mlp_train(pd.data_ptr(), m.data_ptr(), v.data_ptr(), xd.data_ptr(), td.data_ptr(), 1, 1,
                                  256, 1e-3, 1e-5, 0.9, 0.999, 1e-8, 0, 1 << 30, 7, 1, log.data_ptr(),
                                  _stream()) == 0
        pr, _, _ = omlp.adam(p.astype(np.float64), gr, np.zeros_like(gr), np.zeros_like(gr), 1e-3, 0.9, 0.999,
                             1e-8, 1)
        got = pd.cpu().numpy()
        moved = np.abs(gr) > 1e-6        # where |g| >> eps the first step is lr * sign(g)
        np.testing.assert_allclose(got[moved], pr[moved], rtol=0, atol=2e-6)
        assert np.isfinite(float(log[0].item()))

    def test_training_reaches_quarter_sample(self):
        """Training on the VTI model (the last quarter quantisation-aware)
        reaches held-out errors below 0.25 sample (dt 2 ms) under the fused
        inference's bf16 numerics."""
        from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
        from pm_amd.synthetic import vti_model

        vc, ec = vti_model(n_xy=16, n_t=51, x_hi=2000.0, t_hi_ms=2000.0)
        norm = Normalizer(tau_max=1.0, d_max=2000.0, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
        tr = make_training_set(vc, ec, 1_000_000, 2000.0, norm, seed=1)
        ho = make_training_set(vc, ec, 50_000, 2000.0, norm, seed=2)
        s = TraveltimeSurrogate.create(norm, seed=0).train(tr, steps=20000, batch=8192, log_every=2000,
                                                           qat_fraction=0.25)
        err = s.error_samples(ho, 2.0, bf16=True)
        print("loss history", s.history, "held-out error (samples, bf16)", err)
        assert s.history[-1][1] < s.history[0][1] * 1e-2
        assert err["max"] < 0.25


def test_anisotropic_factor_closed_form():
    """c from the exact VTI leg time equals -2 eta s^2 / (1 + 2 eta s^2) with s
    the straight-ray sine (the identity the input encoding rests on), its
    limit at tau = 0, and 0 for vanishing distance; the learned leg time
    rebuilt from c is the exact one."""
    import torch

    from oracle import oracle as orc
    from pm_amd.surrogate import anisotropic_factor, hyperbolic, learned_leg_time

    rng = np.random.default_rng(21)
    tau = rng.uniform(0.01, 2.0, 400)
    d = rng.uniform(1.0, 5000.0, 400)
    v = rng.uniform(1500.0, 4500.0, 400)
    eta = rng.uniform(0.0, 0.2, 400)
    t = np.array([orc.vti_leg(*a) for a in zip(tau, d, v, eta)])
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
    c = anisotropic_factor(T(t), T(tau), T(d), T(v), T(eta)).numpy()
    s = d / (v * hyperbolic(T(tau), T(d), T(v)).numpy())
    np.testing.assert_allclose(c, -2 * eta * s ** 2 / (1 + 2 * eta * s ** 2), rtol=1e-7, atol=1e-10)
    # the learned target g = c / eta: -2 s^2 / (1 + 2 eta s^2), in [-2, 0]
    big = eta > 1e-3
    g = c[big] / eta[big]
    np.testing.assert_allclose(g, -2 * s[big] ** 2 / (1 + 2 * eta[big] * s[big] ** 2), rtol=1e-6, atol=1e-7)
    assert (g <= 0).all() and (g >= -2).all()
    np.testing.assert_allclose(learned_leg_time(T(tau), T(d), T(v), T(c)).numpy(), t, rtol=1e-12)
    c0 = anisotropic_factor(T([0.0]), T([0.0]), T([300.0]), T([2000.0]), T([0.1])).numpy()
    np.testing.assert_allclose(c0, [-0.2 / 1.2])
    cz = anisotropic_factor(T([0.5]), T([0.5]), T([0.0]), T([2000.0]), T([0.1])).numpy()
    assert cz[0] == 0.0


def test_encoding_ranges():
    import torch

    from pm_amd.surrogate import Normalizer

    norm = Normalizer(tau_max=1.0, d_max=2000.0, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    x = torch.tensor([[0.0, 0.0, 2000.0, 1500.0, 0.0], [1.0, 2000.0, 0.0, 4500.0, 0.2],
                      [0.0, 500.0, 0.0, 3000.0, 0.1]], dtype=torch.float64)
    e = norm.apply(x).numpy()
    np.testing.assert_allclose(e[0], [-1, -1, 1, -1, -1], atol=1e-7)   # tau = 0, d > 0: sine 1
    np.testing.assert_allclose(e[1, [0, 3, 4]], [1, 1, 1], atol=1e-7)
    np.testing.assert_allclose(e[2], [-1, 1, -1, 0, 0], atol=1e-7)
    assert np.all(np.abs(e) <= 1 + 1e-6)
