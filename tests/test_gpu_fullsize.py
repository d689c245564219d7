"""GPU checks at the benchmark workloads' full sizes (128^3 hydrogen, 4096
rays; 256^3 H + He, 65,536 rays): the whole 128^3 tick against the oracle,
and at 256^3 the properties that hold at any size (photon conservation over
the tick, bit-exact lists for a sample of rays, the tree's leaf rule)."""
import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd.fields import make_workload
from oracle import pyoracle as po

from conftest import RATES, print_case
from test_gpu_parity import assert_f32_sums, load, oracle_for, run_one_tick

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def ctx():
    c = psr.Context(0)
    yield c
    c.close()


def photons_balance(st, rays):
    """absorbed (first ionizations, in photons) + left in the rays at their
    end = emitted: src/loop.c:479-595 moves each absorbed photon from TM into
    exactly one species' first-ionization count (with secondary ionization
    on and x > 0, which the initial x_HII > 0 gives every hit)"""
    absorbed = st["first_ionization_HI"] + st["first_ionization_HeI"] + st["first_ionization_HeII"]
    emitted = float(rays["Nph"].sum())
    return (absorbed + st["lost_photon_count_sum"]) / emitted - 1.0


def test_full_size_tick_128cube(ctx):
    """workload 2 at its full size: one whole tick (trace, sort, deposit,
    merge, update) against the oracle"""
    print_case("full-size tick 128^3")
    wl = make_workload(2, RATES)
    assert wl.ps.npar == 128 ** 3 and wl.nray == 4096
    orc = oracle_for(wl)
    load(ctx, wl)
    ctx.reset_stats()
    out, ref, rays, (ln, tm, st), off = run_one_tick(ctx, wl, orc, psr.PSR_PHASE_ALL)
    s = ctx.stats()
    assert s["total_deposit_count"] == off[-1]
    assert np.array_equal(out.hits, ref.hits)
    assert np.array_equal(out.lasthit, ref.lasthit)
    # isothermal hydrogen: the closed-form update, same expressions
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6
    assert abs(photons_balance(s, rays)) < 1e-9


def test_full_size_256cube_h_he(ctx):
    """workload 3 at its full size (16.8M particles, 65,536 rays): one trace
    + deposit tick conserves photons; 64 sampled rays' lists equal the
    oracle's bit for bit; leaves hold at most 8 particles"""
    print_case("full-size 256^3 H+He")
    wl = make_workload(3, RATES)
    assert wl.ps.npar == 256 ** 3 and wl.nray == 65536
    load(ctx, wl)
    nn, depth, skipped = ctx.tree_info()
    assert skipped == 0
    e = ctx.tree_export()
    assert e["npar"].max() <= 8
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    ctx.reset_stats()
    ctx.tick(1, wl.tick_time, phases=psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT)
    s = ctx.stats()
    assert s["total_deposit_count"] > 50 * 10 ** 6
    assert abs(photons_balance(s, rays)) < 1e-9
    # deposits are non-negative and land only on hit particles
    dep = wl.ps.copy()
    ctx.state_download(dep)
    assert (dep.yGdepHI >= 0).all() and (dep.heat >= 0).all()
    assert not ((dep.hits == 0) & (dep.yGdepHI > 0)).any()
    # a sample of rays against the oracle's trace on the same field
    pick = np.random.default_rng(7).choice(len(rays), 64, replace=False)
    sample = rays[np.sort(pick)]
    ctx.rays_upload(sample)
    off, ip, d, b = ctx.trace_lists()
    t = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    ooff, oh = po.trace_rays(t, wl.ps, sample)
    assert off[-1] > 64 * 100
    assert np.array_equal(off, ooff)
    assert np.array_equal(ip, oh["ipar"])
    assert np.array_equal(d, oh["d"])
    assert np.array_equal(b, oh["b"])
