"""Ray emission (emit_rays, src/loop.c:163-373): the oracle's restatement
(or_emit_rays) against the host emitter bit for bit and against what the
mathematics fixes (isotropy, draws in proportion to the weights, photon
conservation, ordering, lastemit); the device emission (psr_emit_rays)
against the oracle bit for bit."""
import numpy as np
import pytest
from scipy import stats

import psphray_amd as psr
from psphray_amd.fields import make_workload
from oracle import pyoracle as po

from conftest import RATES, print_case

CASES = [(1, {}), (2, dict(n=16)), (3, dict(n=16)), (5, dict(nfield=12))]


def oracle_emitter(wl):
    em = wl.emitter
    return po.Emitter(em.seed, em.pos, em.ngd, em.field, em.spec, em.spectra, em.boxsizes, em.tick_time)


@pytest.mark.parametrize("cfg,kw", CASES)
def test_oracle_emission_equals_host_emitter(cfg, kw):
    """the C restatement and the numpy emitter draw the same lists from the
    same streams: every byte of every ray, and the same lastemit"""
    wl = make_workload(cfg, RATES, **kw)
    oe = oracle_emitter(wl)
    for t in (1, 2, 3, 7):
        a = wl.rays(t)
        b = oe.emit(t, wl.nray, wl.per_field)
        assert len(a) > 0 and a.tobytes() == b.tobytes(), (cfg, t)
        assert np.array_equal(wl.emitter.lastemit, oe.lastemit)


def test_emission_isotropic_and_unit():
    """directions: unit vectors (to float rounding), z uniform on [-1, 1]
    and the azimuth uniform (Archimedes: isotropy <=> both), mean zero"""
    wl = make_workload(3, RATES, n=8)
    oe = oracle_emitter(wl)
    d = np.concatenate([oe.emit(t, 65536)["dir"].astype(np.float64) for t in (1, 2, 3)])
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-7
    n = len(d)
    assert np.abs(d.mean(0)).max() < 5 / np.sqrt(n)
    assert stats.kstest(d[:, 2], stats.uniform(-1, 2).cdf).pvalue > 1e-3
    phi = np.arctan2(d[:, 1], d[:, 0])
    assert stats.kstest(phi, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-3


def test_emission_draws_sources_by_weight_and_conserves_photons():
    """sources are drawn with probability weight / total (chi-square over
    the 64 sources of workload 2, 40 ticks of fresh sources); a source's
    rays share its weight, so the tick's rays carry the weights of the
    sources that emitted; rays are grouped by source in ray-id order; every
    source that emitted gets lastemit = tick, the others keep theirs and
    carry their weight into the next tick"""
    wl = make_workload(2, RATES, n=8)
    em = wl.emitter
    counts = np.zeros(em.ns)
    expect = np.zeros(em.ns)
    for t in range(1, 41):
        oe = oracle_emitter(wl)        # fresh lastemit: weights Ngamma_dot * t * tick_time
        oe.lastemit[:] = 0
        w = em.ngd * t * em.tick_time
        r = oe.emit(t, 4096)
        src = r["src_or_par"]
        assert (np.diff(src) >= 0).all()
        counts += np.bincount(src, minlength=em.ns)
        expect += 4096 * w / w.sum()
        got = np.zeros(em.ns)
        np.add.at(got, src, r["Nph"])
        hit = np.bincount(src, minlength=em.ns) > 0
        assert np.allclose(got[hit], w[hit], rtol=1e-12) and (got[~hit] == 0).all()
        assert (oe.lastemit[hit] == t).all() and (oe.lastemit[~hit] == 0).all()
    m = expect > 5
    chi2 = ((counts[m] - expect[m]) ** 2 / expect[m]).sum()
    assert stats.chi2.sf(chi2, m.sum() - 1) > 1e-3
    # a source skipped at one tick carries its weight to the next
    oe = oracle_emitter(wl)
    r1 = oe.emit(1, 16)
    quiet = np.setdiff1d(np.arange(em.ns), r1["src_or_par"])
    r2 = oe.emit(2, 4096)
    for s in quiet[:5]:
        sel = r2["src_or_par"] == s
        if sel.any():
            assert r2["Nph"][sel].sum() == pytest.approx(em.ngd[s] * 2 * em.tick_time, rel=1e-12)


def test_emission_spectrum_bins_by_weight():
    """photon energies are the spectrum's bins, drawn in proportion to the
    bin weights (workload 3's 16-bin blackbody, chi-square)"""
    wl = make_workload(3, RATES, n=8)
    oe = oracle_emitter(wl)
    eng, cdf = wl.emitter.spectra[0]
    f = np.concatenate([oe.emit(t, 65536)["freq"] for t in (1, 2)])
    idx = np.searchsorted(eng, f)
    assert np.array_equal(eng[idx], f)
    counts = np.bincount(idx, minlength=len(eng))
    p = np.diff(np.concatenate([[0.0], cdf]))
    expect = len(f) * p
    m = expect > 5
    chi2 = ((counts[m] - expect[m]) ** 2 / expect[m]).sum()
    assert stats.chi2.sf(chi2, m.sum() - 1) > 1e-3


def _device_context(wl):
    ctx = psr.Context(0)
    ctx.set_physics(wl.physics)
    ctx.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)
    ctx.set_sources(wl.emitter)
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,kw", CASES + [(3, {})])
def test_device_emission_equals_oracle(cfg, kw):
    """psr_emit_rays against the oracle: every byte of every ray over four
    ticks (pool offsets, per-field groups), and lastemit"""
    print_case("device emission", cfg, kw)
    wl = make_workload(cfg, RATES, **kw)
    oe = oracle_emitter(wl)
    with _device_context(wl) as ctx:
        first = 0
        for t in (1, 2, 3, 5):
            n = ctx.emit_rays(t, wl.nray, wl.per_field, pool_first=first)
            ref = oe.emit(t, wl.nray, wl.per_field)
            assert n == len(ref) > 0
            got = ctx.rays_download(first, n)
            assert got.tobytes() == ref.tobytes(), (cfg, t)
            first += n
        assert np.array_equal(ctx.sources_lastemit(len(oe.lastemit)), oe.lastemit)


@pytest.mark.gpu
def test_tick_with_device_emitted_rays():
    """a whole tick over rays emitted on the device equals the same tick
    over the same rays built on the host and uploaded, against the oracle"""
    print_case("tick with emitted rays")
    wl = make_workload(2, RATES, n=24, nray=512)
    oe = oracle_emitter(wl)
    rays = oe.emit(1, wl.nray)
    with _device_context(wl) as ctx:
        n = ctx.emit_rays(1, wl.nray)
        assert n == len(rays)
        ctx.tick(1, wl.tick_time, nray=n)
        out = ctx.state_download(wl.ps.copy())
    from test_gpu_parity import oracle_for
    orc = oracle_for(wl)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    rc, _ = po.tick(orc, t, ref, rays, 1, wl.tick_time)
    assert rc == 0
    assert np.array_equal(out.hits, ref.hits)
    assert np.array_equal(out.lasthit, ref.lasthit)
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6


@pytest.mark.gpu
def test_back_to_back_emitted_ticks_equal_oracle():
    """five ticks issued back to back, each over rays the device emits into
    one of three pool ranges: each tick's update runs beside the next
    tick's emission, trace and sort (source rays do not wait for it); the
    state equals five sequential oracle ticks over the same rays (hit
    counts and last hits nearly everywhere, fractions to rounding: the
    closed-form update over deposits summed in another f32 order)"""
    print_case("back-to-back emitted ticks")
    wl = make_workload(2, RATES, n=24, nray=512)
    oe = oracle_emitter(wl)
    room = 2 * wl.nray
    with _device_context(wl) as ctx:
        assert ctx.tuning(psr.PSR_TUNE_TICK_OVERLAP) == 1
        for t in range(1, 6):
            first = (t % 3) * room
            n = ctx.emit_rays(t, wl.nray, pool_first=first)
            assert 0 < n <= room
            ctx.tick(t, wl.tick_time, ray_first=first, nray=n)
        out = ctx.state_download(wl.ps.copy())
        s = ctx.stats()
    from test_gpu_parity import oracle_for
    orc = oracle_for(wl)
    ref = wl.ps.copy()
    tree = po.Tree(ref.pos, ref.sml, wl.boxsize)
    for t in range(1, 6):
        rc, _ = po.tick(orc, tree, ref, oe.emit(t, wl.nray), t, wl.tick_time)
        assert rc == 0
    assert s["total_deposit_count"] > 0
    assert np.mean(out.hits != ref.hits) < 1e-2 and np.mean(out.lasthit != ref.lasthit) < 1e-2
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-5


@pytest.mark.gpu
def test_emission_argument_checks():
    """ticks must increase; a source's field must exist in the tree"""
    print_case("emission checks")
    wl = make_workload(1, RATES, n=8, nray=8)
    with _device_context(wl) as ctx:
        ctx.emit_rays(3, 8)
        with pytest.raises(psr.PsrError):
            ctx.emit_rays(3, 8)
        with pytest.raises(psr.PsrError):
            ctx.emit_rays(2, 8)
        em = wl.emitter
        em.field = np.full(em.ns, 4, np.int32)
        ctx.set_sources(em)
        with pytest.raises(psr.PsrError):
            ctx.emit_rays(4, 8)
