"""Chemistry surrogate: the FP64 oracle network pinned by central
differences and by its definition (CPU), the device forward pass and
gradients against it, and surrogate ticks judged by the exact solver (GPU)."""
import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd import _abi
from psphray_amd.fields import make_workload
from oracle import mlp_ref as mr

from conftest import RATES, print_case


def random_model(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mr.param_count())
    k = 0
    for a, b in mr.SHAPES:
        p[k:k + a * b] = rng.uniform(-1, 1, a * b) * scale * np.sqrt(6.0 / (a + b))
        k += a * b
        p[k:k + b] = rng.normal(0, 0.1, b)
        k += b
    return p


# ---------------------------------------------------------------- CPU pins
def test_oracle_backward_matches_central_differences():
    """the FP64 backward pass against (L(p + h e) - L(p - h e)) / 2h"""
    rng = np.random.default_rng(3)
    p = random_model(3)
    xn = rng.normal(0, 1, (24, mr.IN))
    t = rng.normal(0, 0.3, (24, mr.OUT))
    w = rng.uniform(0.01, 1, (24, mr.OUT))
    _, g = mr.loss_grad(p, xn, t, w)
    idx = np.concatenate([rng.choice(len(p), 200, replace=False),
                          np.arange(len(p) - 4, len(p))])          # output biases too
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mr.loss_grad(p + e, xn, t, w)[0] - mr.loss_grad(p - e, xn, t, w)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k


def test_oracle_forward_matches_layer_formula():
    """one hidden unit by hand: silu(sum_i x_i W1[i][o] + b1[o]) feeds layer 2"""
    p = random_model(5)
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = mr.unpack(p)
    x = np.linspace(-1, 1, mr.IN)[None, :]
    h1 = np.array([[mr.silu(x[0] @ W1[:, o] + b1[o]) for o in range(mr.HID)]])
    h2 = mr.silu(h1 @ W2 + b2)
    h3 = mr.silu(h2 @ W3 + b3)
    assert np.allclose(mr.forward(p, x), h3 @ W4 + b4, rtol=1e-12, atol=1e-12)
    assert mr.param_count() == psr._abi.lib().psr_mlp_param_count() == 10052


def test_features_match_definition():
    """psr_mlp_features (host code of the library) against the definition"""
    rng = np.random.default_rng(11)
    st = np.zeros(8, psr.STEP_DTYPE)
    st["lambdaH"] = rng.uniform(0, 1, 8)
    st["lambdaHeI"] = rng.uniform(0, 0.5, 8)
    st["lambdaHeII"] = rng.uniform(0, 0.5, 8)
    st["T"] = 10 ** rng.uniform(2, 5, 8)
    st["nH"] = 10 ** rng.uniform(-5, 0, 8)
    st["yGdepHI"] = 10 ** rng.uniform(-8, 2, 8)
    st["yGdepHI"][0] = 0.0
    st["heat"] = 10 ** rng.uniform(-20, -10, 8)
    st["yeMET"] = rng.uniform(0, 1e-3, 8)
    st["time"] = 10 ** rng.uniform(-4, 0, 8)
    lib = _abi.lib()
    units = psr.default_units(hmf=0.76)
    got = np.zeros((8, mr.IN), np.float32)
    for k in range(8):
        rc = lib.psr_mlp_features(st[k:k + 1].ctypes.data, C_ref(units), got[k].ctypes.data)
        assert rc == 0
    want = mr.features(st["lambdaH"], st["lambdaHeI"], st["lambdaHeII"], st["T"], st["nH"],
                       st["yGdepHI"], st["yGdepHeI"], st["yGdepHeII"], st["heat"], st["yeMET"],
                       st["time"])
    assert np.allclose(got, want.astype(np.float32), rtol=1e-6, atol=1e-6)
    assert (got[:, 13:21] == 0).all()


def C_ref(u):
    import ctypes
    return ctypes.byref(u)


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def ctx():
    c = psr.Context(0)
    yield c
    c.close()


def set_model(ctx, p, seed=0):
    rng = np.random.default_rng(seed)
    mean = rng.normal(0, 1, mr.IN).astype(np.float32)
    sigma = rng.uniform(0.5, 2, mr.IN).astype(np.float32)
    ctx.mlp_set_params(p.astype(np.float32))
    ctx.mlp_set_norm(mean, sigma)
    return mean.astype(np.float64), sigma.astype(np.float64)


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(ctx):
    """device forward (BF16 MFMA, FP32 accumulation) against the oracle with
    the same BF16 roundings of weights and activations, and against FP64"""
    print_case("mlp-forward")
    p = random_model(7)
    mean, sigma = set_model(ctx, p, 7)
    rng = np.random.default_rng(8)
    n = 1000                                   # not a multiple of 64: a ragged last round
    x = (rng.normal(0, 1, (n, mr.IN)) * sigma + mean).astype(np.float32)
    x[:, 1:4] = rng.uniform(0, 1, (n, 3))      # fractions
    y = ctx.mlp_forward(x)
    xn = (x.astype(np.float64) - mean) / sigma
    out_q = mr.forward(p.astype(np.float32), xn, emulate_bf16=True)
    out = mr.forward(p.astype(np.float32), xn)
    # compare the outputs the predictions came from (logit and log10 T
    # changes); BF16 rounding boundaries can flip with the FP32 exp of SiLU
    got = np.empty_like(out)
    for j in range(3):
        got[:, j] = mr.logit(y[:, j]) - mr.logit(x[:, 1 + j])
    got[:, 3] = y[:, 3].astype(np.float64) - x[:, 4]
    scale = np.abs(out).max()
    ok = (np.abs(y[:, :3] - 0.5) < 0.49).all(1)  # logit well conditioned (float32 results)
    assert ok.sum() > n // 2
    assert np.abs(got - out_q)[ok].max() < 2e-2 * scale
    assert np.median(np.abs(got - out_q)[ok]) < 2e-3 * scale
    assert np.abs(got - out)[ok].max() < 6e-2 * scale
    assert np.allclose(y, mr.predict(x, out_q), atol=2e-2 * scale)


@pytest.mark.gpu
def test_mlp_gradient_matches_fp64_oracle(ctx):
    print_case("mlp-grad")
    p = random_model(9, scale=0.7)
    mean, sigma = set_model(ctx, p, 9)
    rng = np.random.default_rng(10)
    n = 777
    x = rng.normal(0, 1, (n, mr.IN)) * sigma + mean
    x = x.astype(np.float32)
    x[:, 1:4] = rng.uniform(0, 1, (n, 3))
    y = np.concatenate([rng.uniform(0, 1, (n, 3)), x[:, 4:5] + rng.normal(0, 0.5, (n, 1))], 1)
    y = y.astype(np.float32)
    g, loss = ctx.mlp_grad(x, y)
    xn = (x.astype(np.float64) - mean) / sigma
    tt, ww = mr.target(x, y)
    lref, gref = mr.loss_grad(p.astype(np.float32), xn, tt, ww)
    assert abs(loss - lref) < 2e-2 * lref
    cos = g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert cos > 0.999
    assert np.linalg.norm(g - gref) < 4e-2 * np.linalg.norm(gref)
    # per layer, so that no block is wrong while the total looks fine
    k = 0
    for a, b in mr.SHAPES:
        for m in (a * b, b):
            blk = slice(k, k + m)
            assert np.linalg.norm(g[blk] - gref[blk]) < 6e-2 * np.linalg.norm(gref[blk]) + 1e-7, (a, b, m)
            k += m


@pytest.mark.gpu
def test_mlp_training_fits_a_smooth_map():
    """training on the device lowers the loss of a known smooth target by
    two orders of magnitude"""
    print_case("mlp-train")
    rng = np.random.default_rng(12)
    n = 20000
    x = rng.normal(0, 1, (n, mr.IN)).astype(np.float32)
    x[:, 1:4] = rng.uniform(0.05, 0.95, (n, 3))
    x[:, 10:13] = mr.logit(x[:, 1:4])
    x[:, 13:21] = 0
    out = np.stack([0.8 * np.tanh(x[:, 0] + 0.5 * x[:, 5]), 0.3 * x[:, 2] * x[:, 3],
                    0.5 * np.sin(x[:, 21]), 0.05 * x[:, 4]], 1)
    y = mr.predict(x, out).astype(np.float32)
    t, w = mr.target(x, y)
    l0 = float(((w * t) ** 2).sum(1).mean())
    with psr.Context(0) as c:
        loss = c.mlp_train(x, y, epochs=30, batch=1024, lr=3e-3, seed=5)
        pred = c.mlp_forward(x[:2000])
    err = pred - y[:2000]
    print("loss", l0, "->", loss, "max err", np.abs(err).max())
    assert loss < 1e-2 * l0
    assert np.abs(err[:, :3]).mean() < 5e-3 and np.abs(err[:, :3]).max() < 6e-2


def _ticks(ctx, wl, firsts, counts, t0, t1, chem):
    for t in range(t0, t1 + 1):
        ctx.tick(t, wl.tick_time, ray_first=int(firsts[t - 1]), nray=int(counts[t - 1]), chem=chem)


@pytest.mark.gpu
def test_surrogate_tick_against_exact_solver():
    """the exact solver writes training tuples over 60 ticks of a small
    workload-2 field; the network trained on them then runs ticks 61..64
    (falling back to the exact solver where its check fails) from the same
    state as exact ticks 61..64, with the ionisation-balance check (the
    default): ionized fractions agree within 1e-3.  With the conservation
    residual alone nearly every prediction is taken; the test records its
    error, which the balance check exists to bound."""
    print_case("surrogate-tick")
    wl = make_workload(2, RATES, n=24, nray=256)
    T, K = 60, 4
    ticks = [wl.rays(t) for t in range(1, T + K + 1)]
    counts = np.array([len(r) for r in ticks])
    firsts = np.concatenate([[0], np.cumsum(counts)])
    pool = np.concatenate(ticks)
    res = {}
    with psr.Context(0) as c:
        c.set_physics(wl.physics)
        c.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)
        c.rays_upload(pool)
        c.mlp_capture(2_000_000)
        _ticks(c, wl, firsts, counts, 1, T, psr.PSR_CHEM_EXACT)
        n, _, _ = c.mlp_captured()
        print("captured", n)
        assert n > 50000
        loss = c.mlp_train(n=0, epochs=40, batch=4096, lr=2e-3, seed=3)
        print("train loss", loss)
        c.mlp_capture(0)                       # the judged ticks are not trained on
        snap = wl.ps.copy()
        c.state_download(snap)
        _ticks(c, wl, firsts, counts, T + 1, T + K, psr.PSR_CHEM_EXACT)
        ex = wl.ps.copy()
        c.state_download(ex)
        for balance in (False, True):
            c.mlp_set_balance_check(balance)
            c.state_upload(snap)
            c.reset_stats()
            _ticks(c, wl, firsts, counts, T + 1, T + K, psr.PSR_CHEM_SURROGATE)
            sg = wl.ps.copy()
            c.state_download(sg)
            res[balance] = (sg, c.stats())
    for balance, (sg, st) in res.items():
        acc, fb = st["surrogate_count"], st["surrogate_fallbacks"]
        dx = np.abs(sg.lambdaH - ex.lambdaH)
        print("balance", balance, "surrogate", acc, "fallbacks", fb,
              "x_HII error max %.2e p99 %.2e median %.2e" % (dx.max(), np.percentile(dx, 99), np.median(dx)))
        assert np.array_equal(sg.lasthit, ex.lasthit)
        if balance:
            assert acc > 0.5 * (acc + fb)
            assert dx.max() <= 1e-3
        else:
            assert acc > 0.95 * (acc + fb)
            assert np.median(dx) <= 1e-3


@pytest.mark.gpu
def test_surrogate_tick_off_the_spot_runs_the_exact_solver():
    """the network predicts no recombination counters, so a run without
    on-the-spot recombination takes the exact solver in surrogate ticks: no
    update is counted as the network's, and the state equals an exact
    tick's up to the f32 summation order of the two runs' deposits (which
    an adaptive step decision can carry to the solver's tolerance)"""
    print_case("surrogate off-the-spot")
    wl = make_workload(1, RATES, n=16, nray=64)
    wl.physics = psr.Physics(RATES, wl.physics.units,
                             psr.make_flags(isothermal=True, h_only=True, on_the_spot=False))
    rays = wl.rays(1)
    out = {}
    for chem in (psr.PSR_CHEM_EXACT, psr.PSR_CHEM_SURROGATE):
        with psr.Context(0) as c:
            c.set_physics(wl.physics)
            c.load(wl.ps, boxsize=wl.boxsize)
            rng = np.random.default_rng(2)
            c.mlp_set_params(rng.normal(0, 0.05, _abi.lib().psr_mlp_param_count()).astype(np.float32))
            c.mlp_set_norm(np.zeros(22), np.ones(22), np.full(22, -1e30), np.full(22, 1e30))
            c.rays_upload(rays)
            c.tick(1, wl.tick_time, chem=chem)
            st = wl.ps.copy()
            c.state_download(st)
            out[chem] = (st, c.stats())
    (a, sa), (b, sb) = out[psr.PSR_CHEM_EXACT], out[psr.PSR_CHEM_SURROGATE]
    assert sb["surrogate_count"] == 0 and sb["active_count"] == sa["active_count"] > 0
    assert np.array_equal(a.hits, b.hits) and np.array_equal(a.lasthit, b.lasthit)
    for f in ("lambdaH", "yGrecHII"):
        va, vb = getattr(a, f).astype(np.float64), getattr(b, f).astype(np.float64)
        d = np.abs(va - vb)
        assert np.median(d) < 1e-9 and (d <= 1e-3 * np.maximum(np.abs(va), np.abs(vb)) + 1e-6).all(), f


@pytest.mark.gpu
def test_surrogate_stiff_scope_against_exact_solver():
    """PSR_CHEM_SURROGATE_STIFF on an H + He field with temperature: the
    network sees only the particles the exact solver's first attempt leaves
    unfinished (its accepted and rejected counts add up to them; a rejected
    one continues from the saved integrator state), and the ticks agree with
    exact ticks from the same state within 1e-3 in x_HII; on the closed-form
    hydrogen field nothing is left for the network"""
    print_case("surrogate-stiff")
    wl = make_workload(3, RATES, n=16, nray=256)
    T, K = 24, 3
    ticks = [wl.rays(t) for t in range(1, T + K + 1)]
    counts = np.array([len(r) for r in ticks])
    firsts = np.concatenate([[0], np.cumsum(counts)])
    pool = np.concatenate(ticks)
    with psr.Context(0) as c:
        c.set_physics(wl.physics)
        c.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)
        c.rays_upload(pool)
        c.mlp_capture(2_000_000)
        _ticks(c, wl, firsts, counts, 1, T, psr.PSR_CHEM_EXACT)
        n, _, _ = c.mlp_captured()
        assert n > 20000
        c.mlp_train(n=0, epochs=30, batch=4096, lr=2e-3, seed=5)
        c.mlp_capture(0)
        snap = wl.ps.copy()
        c.state_download(snap)
        _ticks(c, wl, firsts, counts, T + 1, T + K, psr.PSR_CHEM_EXACT)
        ex = wl.ps.copy()
        c.state_download(ex)
        c.state_upload(snap)
        c.reset_stats()
        _ticks(c, wl, firsts, counts, T + 1, T + K, psr.PSR_CHEM_SURROGATE_STIFF)
        sg = wl.ps.copy()
        c.state_download(sg)
        st = c.stats()
    acc, fb = st["surrogate_count"], st["surrogate_fallbacks"]
    print("stiff: surrogate", acc, "fallbacks", fb, "active", st["active_count"])
    assert acc > 0 and acc + fb < st["active_count"]
    dx = np.abs(sg.lambdaH - ex.lambdaH)
    assert dx.max() < 1e-3
    # the closed form finishes everything in phase 1
    wl2 = make_workload(2, RATES, n=12, nray=64)
    with psr.Context(0) as c:
        c.set_physics(wl2.physics)
        c.load(wl2.ps, wl2.field_offsets, wl2.boxsizes, boxsize=wl2.boxsize)
        rays = wl2.rays(1)
        c.rays_upload(rays)
        c.mlp_capture(100000)
        c.tick(1, wl2.tick_time)
        c.sync()
        c.mlp_train(n=0, epochs=2, batch=256, lr=2e-3, seed=5)
        c.mlp_capture(0)
        c.reset_stats()
        c.tick(2, wl2.tick_time, chem=psr.PSR_CHEM_SURROGATE_STIFF)
        c.sync()
        st = c.stats()
    assert st["surrogate_count"] == 0 and st["surrogate_fallbacks"] == 0 and st["active_count"] > 0
