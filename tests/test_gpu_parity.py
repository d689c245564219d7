"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer/index work is compared bit-exactly, floating
point within the tolerances stated per test."""
import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd.fields import make_workload
from oracle import pyoracle as po

from conftest import RATES, print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = psr.Context(0)
    yield c
    c.close()


def oracle_for(wl):
    u = wl.physics.units
    f = wl.physics.flags
    ou = po.units(hmf=u.C_HMF, hemf=u.C_HEMF)
    of = po.OrFlags(f.isothermal, f.adiabatic, f.on_the_spot, f.h_only, f.secondary_ionization,
                    f.trace_only, f.no_photon, f.comoving, f.redshift, f.fake_temperature)
    # the oracle's own copy of src/kernel.akline (the product loads the same
    # file as data, psr_tables_load_kernel)
    return po.Oracle(ou, of, rates=RATES, kernel=po.reference_kernel_table())


def load(ctx, wl):
    ctx.set_physics(wl.physics)
    ctx.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)


def oracle_lists(wl, rays):
    """per-field oracle trees; lists in the caller's particle indices"""
    ps = wl.ps
    if wl.field_offsets is None:
        t = po.Tree(ps.pos, ps.sml, wl.boxsize)
        return po.trace_rays(t, ps, rays)
    offs = [0]
    out = []
    for i in range(len(rays)):
        f = int(rays["field"][i])
        a, b = wl.field_offsets[f], wl.field_offsets[f + 1]
        sub = ps.copy()
        for name in sub.fields():
            setattr(sub, name, np.ascontiguousarray(getattr(ps, name)[a:b]))
        sub.npar = b - a
        t = po.Tree(sub.pos, sub.sml, wl.boxsizes[f])
        o, h = po.trace_rays(t, sub, rays[i:i + 1], nthreads=1)
        h = h.copy()
        h["ipar"] += a
        out.append(h)
        offs.append(offs[-1] + len(h))
    return np.array(offs, np.int64), (np.concatenate(out) if out else np.zeros(0, po.HIT_DTYPE))


def assert_f32_sums(a, r, hits, name):
    """f32 accumulators summed in another order (atomics, or LDS sums of
    per-hit records).  Both sides add positive terms (at most 4 per hit:
    the reference's separate increments), so each is within 4 * hits * 2^-24
    of the exact sum (recursive summation bound); 1e-5 rel covers the rest"""
    a, r = a.astype(np.float64), r.astype(np.float64)
    tol = (1e-5 + 8.0 * hits * 2.0 ** -24) * np.abs(r) + 1e-6 * max(np.abs(r).max(), 1e-300)
    bad = np.abs(a - r) > tol
    assert not bad.any(), (name, int(bad.sum()), a[bad][:4], r[bad][:4], hits[bad][:4])


CASES = [(1, 16, 64), (1, 32, 32), (2, 24, 256), (3, 20, 256)]
# deposit modes: one ray per workgroup with atomics or binned records, and
# four rays per workgroup (atomics)
ACCUM = [(psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_SINGLE), (psr.PSR_ACCUM_BINNED, psr.PSR_WALK_SINGLE),
         (psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_MULTI), (psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_MULTI2),
         (psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_MULTI8)]
ACCUM_IDS = ["atomic", "binned", "multi", "multi2", "multi8"]


@pytest.fixture
def accum(ctx, request):
    """the deposit's mode, fixed for the test, automatic afterwards"""
    a, w = request.param
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, psr.PSR_ACCUM_AUTO)
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_WALK, w)
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, a)
    yield request.param
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, psr.PSR_ACCUM_AUTO)
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_WALK, psr.PSR_WALK_AUTO)


@pytest.mark.parametrize("cfg,n,nray", CASES)
def test_tree_matches_oracle(ctx, cfg, n, nray):
    print_case("tree", cfg, n)
    wl = make_workload(cfg, RATES, n=n, nray=nray)
    load(ctx, wl)
    e = ctx.tree_export()
    t = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    eo = t.export()
    assert len(e["npar"]) == len(eo["npar"])
    # same nodes: multiset of (box, leaf size) bit-exact
    dev = sorted(zip(map(tuple, e["bot"].view(np.int32)), map(tuple, e["top"].view(np.int32)),
                     e["npar"].tolist()))
    orc = sorted(zip(map(tuple, eo["bot"].view(np.int32)), map(tuple, eo["top"].view(np.int32)),
                     [(-1 if fc != -1 else int(np_)) for fc, np_ in zip(eo["first_child"], eo["npar"])]))
    assert dev == orc
    nn, depth, skipped = ctx.tree_info()
    assert depth == eo["depth"].max() and skipped == 0


@pytest.fixture(params=[256, 128], ids=["nt256", "nt128"])
def trace_threads(ctx, request):
    """the traversal's threads per ray, fixed for the test, timed afterwards"""
    ctx.set_tuning(psr.PSR_TUNE_TRACE_THREADS, request.param)
    yield request.param
    ctx.set_tuning(psr.PSR_TUNE_TRACE_THREADS, -1)


@pytest.mark.parametrize("cfg,n,nray", CASES)
def test_trace_lists_bit_exact(ctx, trace_threads, cfg, n, nray):
    print_case("trace", cfg, n, nray, trace_threads)
    wl = make_workload(cfg, RATES, n=n, nray=nray)
    load(ctx, wl)
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    ooff, oh = oracle_lists(wl, rays)
    assert off[-1] > 0
    assert np.array_equal(off, ooff)
    assert np.array_equal(ip, oh["ipar"])
    assert np.array_equal(d, oh["d"])
    assert np.array_equal(b, oh["b"])


def long_ray_field():
    """a 64^3 lattice with smoothing lengths tripled and diagonal rays through
    it: about 11,800 hits per ray, past the LDS sort tiers (8192), beside
    the workload's ordinary rays"""
    wl = make_workload(1, RATES, n=64, nray=24)
    wl.ps.sml[:] *= np.float32(3.0)
    box = wl.boxsize
    r = wl.rays(1)
    rng = np.random.default_rng(7)
    d = np.array([1.0, 1.0, 1.0]) + rng.normal(0, 0.02, (8, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r[:8]["s"] = np.float32(0.01 * box)
    r[:8]["dir"] = d.astype(np.float32)
    r[:8]["length"] = box * np.sqrt(3.0)
    return wl, r


def test_lists_longer_than_the_lds_sort(ctx):
    """lists of more than 8192 hits take the global-memory bitonic sort:
    bit-exact lists, then one whole tick against the oracle"""
    print_case("long lists")
    wl, rays = long_ray_field()
    load(ctx, wl)
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    t = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    ooff, oh = po.trace_rays(t, wl.ps, rays)
    assert np.diff(ooff).max() > 8192 and np.diff(ooff).min() < 4096
    assert np.array_equal(off, ooff)
    assert np.array_equal(ip, oh["ipar"])
    assert np.array_equal(d, oh["d"])
    assert np.array_equal(b, oh["b"])
    orc = oracle_for(wl)
    ctx.reset_stats()
    ctx.tick(1, wl.tick_time)
    out = wl.ps.copy()
    ctx.state_download(out)
    ref = wl.ps.copy()
    po.deposit(orc, ref, rays, ooff, oh)
    act = po.merge(ref, ooff, oh)
    rc, _ = po.update(orc, ref, act, 1, wl.tick_time)
    assert rc == 0
    assert ctx.stats()["total_deposit_count"] == ooff[-1]
    assert np.array_equal(out.hits, ref.hits)
    assert np.array_equal(out.lasthit, ref.lasthit)
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6


def test_trace_batched_fields(ctx, trace_threads):
    print_case("trace batched", trace_threads)
    wl = make_workload(5, RATES, nfield=6, nray=64)
    load(ctx, wl)
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    ooff, oh = oracle_lists(wl, rays)
    assert np.array_equal(off, ooff)
    assert np.array_equal(ip, oh["ipar"])
    assert np.array_equal(d, oh["d"])


def _field_view(ps, a, b):
    sub = ps.copy()
    for name in sub.fields():
        setattr(sub, name, np.ascontiguousarray(getattr(ps, name)[a:b]))
    sub.npar = b - a
    return sub


def test_batched_fields_whole_ticks_against_per_field_oracle(ctx):
    """batched uneven fields (workload 5 at 12 fields) over three whole
    ticks (trace, deposit, merge, update) in one launch per phase, against
    the oracle run field by field on each field's own rays and tree:
    hits and last hits bit-exact, x_HI to 1e-6 (isothermal hydrogen, the
    closed-form update), deposits of the last tick within the f32 summation
    bound"""
    print_case("batched whole ticks")
    wl = make_workload(5, RATES, nfield=12, nray=128)
    orc = oracle_for(wl)
    load(ctx, wl)
    fo = np.asarray(wl.field_offsets, np.int64)
    refs = [_field_view(wl.ps, fo[f], fo[f + 1]) for f in range(len(fo) - 1)]
    trees = [po.Tree(r.pos, r.sml, wl.boxsizes[f]) for f, r in enumerate(refs)]
    for tick in (1, 2, 3):
        rays = wl.rays(tick)
        ctx.rays_upload(rays)
        ctx.tick(tick, wl.tick_time, phases=psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT
                 if tick == 3 else psr.PSR_PHASE_ALL)
        for f, ref in enumerate(refs):
            fr = rays[rays["field"] == f]
            if tick == 3:
                off, hits = po.trace_rays(trees[f], ref, fr)
                po.deposit(orc, ref, fr, off, hits)
            else:
                rc, _ = po.tick(orc, trees[f], ref, fr, tick, wl.tick_time)
                assert rc == 0
    out = ctx.state_download(wl.ps.copy())
    for f, ref in enumerate(refs):
        a, b = fo[f], fo[f + 1]
        sub = _field_view(out, a, b)
        assert np.array_equal(sub.hits, ref.hits), f
        assert np.array_equal(sub.lasthit, ref.lasthit), f
        assert np.abs(sub.lambdaH - ref.lambdaH).max() < 1e-6, f
        for name in ("yGdepHI", "heat"):
            assert_f32_sums(getattr(sub, name), getattr(ref, name), ref.hits, name)
    assert (out.hits > 0).sum() > 1000


def test_empty_field_and_empty_tick(ctx):
    """batched fields with one field emptied: its rays find nothing, the
    others' lists are the oracle's; a tick with no rays leaves the state as
    it was; a whole tick then counts every listed hit"""
    print_case("empty field / empty tick")
    wl = make_workload(5, RATES, nfield=4, nray=32)
    fo = np.asarray(wl.field_offsets, np.int64)
    keep = np.ones(wl.ps.npar, bool)
    keep[fo[1]:fo[2]] = False
    ps = psr.ParticleSystem(int(keep.sum()))
    for f in ps.fields():
        setattr(ps, f, np.ascontiguousarray(getattr(wl.ps, f)[keep]))
    fo2 = fo.copy()
    fo2[2:] -= fo[2] - fo[1]
    wl.ps, wl.field_offsets = ps, fo2
    load(ctx, wl)
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    empty = rays["field"] == 1
    assert empty.any()
    assert (np.diff(off)[empty] == 0).all()
    ooff, oh = oracle_lists(wl, rays[~empty])
    assert off[-1] > 0 and np.array_equal(np.diff(off)[~empty], np.diff(ooff))
    assert np.array_equal(ip, oh["ipar"]) and np.array_equal(d, oh["d"]) and np.array_equal(b, oh["b"])
    before = wl.ps.copy()
    ctx.state_download(before)
    ctx.tick(1, wl.tick_time, ray_first=0, nray=0)
    after = wl.ps.copy()
    ctx.state_download(after)
    for f in before.fields():
        assert np.array_equal(getattr(before, f), getattr(after, f)), f
    ctx.reset_stats()
    ctx.tick(2, wl.tick_time)
    assert ctx.stats()["total_deposit_count"] == off[-1]


def test_trace_only_tick(ctx):
    """a tick with the trace phase alone (traceOnly, src/loop.c:131-148):
    the hits are counted, the particles are left as they were and the
    active set is dropped, so a whole tick afterwards equals one from the
    start"""
    print_case("trace-only tick")
    wl = make_workload(2, RATES, n=24, nray=256)
    orc = oracle_for(wl)
    load(ctx, wl)
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    ctx.reset_stats()
    ctx.tick(1, wl.tick_time, phases=psr.PSR_PHASE_TRACE)
    s = ctx.stats()
    t = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    ooff, _ = po.trace_rays(t, wl.ps, rays)
    assert s["total_deposit_count"] == ooff[-1] and s["active_count"] == 0
    mid = wl.ps.copy()
    ctx.state_download(mid)
    for f in mid.fields():
        assert np.array_equal(getattr(mid, f), getattr(wl.ps, f)), f
    out, ref, _, _, _ = run_one_tick(ctx, wl, orc, psr.PSR_PHASE_ALL)
    assert np.array_equal(out.lasthit, ref.lasthit)
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6


def test_trace_invalid_and_outside_particles(ctx):
    print_case("trace flags")
    wl = make_workload(1, RATES, n=16, nray=64)
    ps = wl.ps
    ps.flag[::7] = 1                       # PF_INVALID: never recorded
    ps.pos[5] = [wl.boxsize * 1.5, 0.1, 0.1]   # outside the root: skipped
    ps.pos[9] = [-0.01, 0.2, 0.2]
    load(ctx, wl)
    _, _, skipped = ctx.tree_info()
    assert skipped == 2
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    t = po.Tree(ps.pos, ps.sml, wl.boxsize)
    ooff, oh = po.trace_rays(t, ps, rays)
    assert np.array_equal(off, ooff) and np.array_equal(ip, oh["ipar"])
    assert not np.isin(ip, np.r_[np.arange(0, ps.npar, 7), 5, 9]).any()


def test_trace_edge_rays(ctx):
    print_case("trace edge rays")
    wl = make_workload(1, RATES, n=16, nray=8)
    load(ctx, wl)
    rays = wl.rays(1)
    rays["length"][0] = 0.0                          # zero-length segment
    rays["s"][1] = [-5.0, -5.0, -5.0]                # starts far outside
    rays["dir"][1] = [-0.57735026, -0.57735026, -0.57735026]
    rays["dir"][2] = [1.0, 0.0, 0.0]                 # axis-aligned
    ctx.rays_upload(rays)
    off, ip, d, b = ctx.trace_lists()
    t = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    ooff, oh = po.trace_rays(t, wl.ps, rays)
    assert np.array_equal(off, ooff) and np.array_equal(ip, oh["ipar"])
    assert off[2] - off[1] == 0


def run_one_tick(ctx, wl, orc, phases, tick=1):
    rays = wl.rays(tick)
    ctx.rays_upload(rays)
    ctx.tick(tick, wl.tick_time, phases=phases)
    out = wl.ps.copy()
    ctx.state_download(out)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    off, hits = po.trace_rays(t, ref, rays)
    ln, tm, st = po.deposit(orc, ref, rays, off, hits)
    if phases & psr.PSR_PHASE_UPDATE:
        act = po.merge(ref, off, hits)
        rc, st2 = po.update(orc, ref, act, tick, wl.tick_time)
        assert rc == 0
    return out, ref, rays, (ln, tm, st), off


@pytest.mark.parametrize("accum", ACCUM, indirect=True, ids=ACCUM_IDS)
@pytest.mark.parametrize("cfg,n,nray", [(1, 32, 32), (2, 24, 256), (3, 20, 128)])
def test_deposit_parity(ctx, cfg, n, nray, accum):
    print_case("deposit", cfg, n, nray, accum)
    wl = make_workload(cfg, RATES, n=n, nray=nray)
    orc = oracle_for(wl)
    load(ctx, wl)
    ctx.reset_stats()
    out, ref, rays, (ln, tm, st), off = run_one_tick(ctx, wl, orc,
                                                     psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT)
    assert np.array_equal(out.hits, ref.hits)
    for f in ("yGdepHI", "yGdepHeI", "yGdepHeII", "heat"):
        assert_f32_sums(getattr(out, f), getattr(ref, f), ref.hits, f)
    dln, dtm = ctx.ray_results(len(rays))
    assert np.array_equal(dln, ln)
    assert np.allclose(dtm, tm, rtol=1e-9, atol=1e-12 * rays["Nph"].max())
    s = ctx.stats()
    assert s["total_deposit_count"] == st.total_deposit_count == off[-1]
    assert s["saturated_deposit_count"] == st.saturated_deposit_count


@pytest.mark.parametrize("accum", ACCUM, indirect=True, ids=ACCUM_IDS)
@pytest.mark.parametrize("he", [False, True])
def test_deposit_no_photon_and_cutoff(ctx, he, accum):
    """rays of negative type start with TM = 0 under no_photon (src/loop.c:411);
    the first live hit then ends the ray; other rays run into the cut-off"""
    print_case("deposit-no-photon", he, accum)
    wl = make_workload(3 if he else 1, RATES, n=16 if he else 24, nray=64)
    f = wl.physics.flags
    flags = psr.make_flags(isothermal=bool(f.isothermal), on_the_spot=bool(f.on_the_spot),
                           h_only=bool(f.h_only), secondary_ionization=bool(f.secondary_ionization),
                           no_photon=True)
    wl.physics = psr.Physics(RATES, wl.physics.units, flags, with_secion=he)
    orc = oracle_for(wl)
    load(ctx, wl)
    ctx.reset_stats()
    rays = wl.rays(1)
    rays["type"][::3] = -1
    rays["Nph"][1::3] *= 1e-6                         # weak rays: cut-off early
    ctx.rays_upload(rays)
    ctx.tick(1, wl.tick_time, phases=psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT)
    out = wl.ps.copy()
    ctx.state_download(out)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    off, hits = po.trace_rays(t, ref, rays)
    ln, tm, st = po.deposit(orc, ref, rays, off, hits)
    assert np.array_equal(out.hits, ref.hits)
    for fld in ("yGdepHI", "yGdepHeI", "yGdepHeII", "heat"):
        assert_f32_sums(getattr(out, fld), getattr(ref, fld), ref.hits, fld)
    dln, dtm = ctx.ray_results(len(rays))
    assert np.array_equal(dln, ln)
    assert np.allclose(dtm, tm, rtol=1e-9, atol=1e-12 * rays["Nph"].max())
    s = ctx.stats()
    assert s["deposit_processed"] < s["total_deposit_count"]   # some rays were cut
    assert s["saturated_deposit_count"] == st.saturated_deposit_count


@pytest.mark.parametrize("he", [False, True])
def test_binned_deposit_full_bins(he):
    """rays from the corner all cross the lower octant, which is the first two
    particle bins (2048 particles each, tree order); their even first share of
    the record buffer (16 per particle + 2^20 records over the 16 bins)
    overflows, and those records go straight to the particles.  The second
    tick runs with rooms planned from the first tick's claims.  Both ticks'
    sums equal the oracle's."""
    print_case("binned-full", he)
    wl = make_workload(3 if he else 1, RATES, n=32, nray=8192)
    orc = oracle_for(wl)
    rays = wl.rays(1)
    rays["s"] = np.float32(0.0)
    # a context of its own: the bins are sized at its first binned deposit
    with psr.Context(0) as c2:
        load(c2, wl)
        c2.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, psr.PSR_ACCUM_BINNED)
        c2.rays_upload(rays)
        for t in (1, 2):
            c2.tick(t, wl.tick_time, phases=psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT)
        out = wl.ps.copy()
        c2.state_download(out)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    off, hits = po.trace_rays(t, ref, rays)
    nb = (wl.ps.npar + 2047) // 2048
    room = (16 * wl.ps.npar + 2 ** 20) // nb
    octant = (wl.ps.pos < wl.boxsize / 2).all(1)
    assert octant.sum() == 2 * 2048
    slot_hits = np.bincount(hits["ipar"], minlength=wl.ps.npar)
    assert slot_hits[octant].sum() > 2 * room      # the case this test is for
    po.deposit(orc, ref, rays, off, hits)
    assert np.array_equal(out.hits, 2 * ref.hits)
    for fld in ("yGdepHI", "yGdepHeI", "yGdepHeII", "heat"):
        assert_f32_sums(getattr(out, fld), 2 * getattr(ref, fld).astype(np.float64), 2 * ref.hits, fld)


def test_deposit_accumulation_is_tuned(ctx):
    """automatic deposit mode: the first ten deposit ticks cycle through
    one-ray atomic and binned sums and two, four and eight rays per 4-wave
    workgroup (each timed twice); a choice is made once they are done, and
    every tick's sums are right whichever ran"""
    print_case("autotune")
    wl = make_workload(2, RATES, n=24, nray=256)
    orc = oracle_for(wl)
    load(ctx, wl)
    assert ctx.tuning() == psr.PSR_ACCUM_AUTO
    ctx.reset_stats()
    rays = wl.rays(1)
    ctx.rays_upload(rays)
    nt = 10
    for t in range(1, nt + 1):
        ctx.tick(t, wl.tick_time, phases=psr.PSR_PHASE_TRACE | psr.PSR_PHASE_DEPOSIT)
    ctx.sync()
    assert ctx.tuning() in (psr.PSR_ACCUM_ATOMIC, psr.PSR_ACCUM_BINNED)
    assert ctx.tuning(psr.PSR_TUNE_DEPOSIT_WALK) in (psr.PSR_WALK_SINGLE, psr.PSR_WALK_MULTI,
                                                     psr.PSR_WALK_MULTI2, psr.PSR_WALK_MULTI8)
    out = wl.ps.copy()
    ctx.state_download(out)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    off, hits = po.trace_rays(t, ref, rays)
    po.deposit(orc, ref, rays, off, hits)
    assert np.array_equal(out.hits, nt * ref.hits)
    for fld in ("yGdepHI", "heat"):
        assert_f32_sums(getattr(out, fld), nt * getattr(ref, fld).astype(np.float64), nt * ref.hits, fld)


@pytest.mark.parametrize("accum,walk,overlap", [(psr.PSR_ACCUM_AUTO, psr.PSR_WALK_AUTO, 1),
                                                (psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_SINGLE, 1),
                                                (psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_MULTI, 1),
                                                (psr.PSR_ACCUM_AUTO, psr.PSR_WALK_AUTO, 0)])
def test_overflowed_tick_is_replayed(accum, walk, overlap):
    """ticks are launched without waiting for the trace's capacity check; a
    tick whose lists outgrow the capacity (2048 hits per ray at first) stops
    every later kernel and is replayed, with the ticks after it, at the next
    call -- the result equals three sequential oracle ticks.  With overlap
    the second tick's trace overflows while the first tick's update runs
    beside it: that update still completes, and only the ticks from the
    second on are replayed.  With a fixed accumulation the ticks are the
    tick-graph trials (kernel by kernel, then a graph launch)."""
    print_case("overflow-replay", accum, walk, overlap)
    wl = make_workload(1, RATES, n=104, nray=16)
    box = wl.boxsize
    r1, r2, r3 = wl.rays(1), wl.rays(2), wl.rays(3)
    rng = np.random.default_rng(7)
    d = np.array([1.0, 1.0, 1.0]) + rng.normal(0, 0.02, (len(r2), 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r2["s"] = np.float32(0.01 * box)
    r2["dir"] = d.astype(np.float32)
    r2["length"] = box * np.sqrt(3.0)
    pool = np.concatenate([r1, r2, r3])
    orc = oracle_for(wl)
    with psr.Context(0) as c2:
        c2.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, accum)
        c2.set_tuning(psr.PSR_TUNE_DEPOSIT_WALK, walk)
        c2.set_tuning(psr.PSR_TUNE_TICK_OVERLAP, overlap)
        load(c2, wl)
        c2.rays_upload(pool)
        for t, (a, rr) in enumerate(((0, r1), (len(r1), r2), (len(r1) + len(r2), r3)), start=1):
            c2.tick(t, wl.tick_time, ray_first=a, nray=len(rr))
        out = wl.ps.copy()
        c2.state_download(out)
        s = c2.stats()
    ref = wl.ps.copy()
    tree = po.Tree(ref.pos, ref.sml, box)
    total = 0
    for t, rr in enumerate((r1, r2, r3), start=1):
        off, _ = po.trace_rays(tree, ref, rr)
        if t == 2:
            assert np.diff(off).max() > 2048          # the case this test is for
        total += int(off[-1])
        rc, _ = po.tick(orc, tree, ref, rr, t, wl.tick_time)
        assert rc == 0
    assert s["total_deposit_count"] == total
    assert np.array_equal(out.lasthit, ref.lasthit)
    assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6


@pytest.mark.parametrize("cfg,accum,walk", [(2, psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_SINGLE),
                                            (2, psr.PSR_ACCUM_BINNED, psr.PSR_WALK_SINGLE),
                                            (3, psr.PSR_ACCUM_BINNED, psr.PSR_WALK_SINGLE),
                                            (3, psr.PSR_ACCUM_ATOMIC, psr.PSR_WALK_MULTI)])
def test_tick_graph_matches_kernel_by_kernel(cfg, accum, walk):
    """ten ticks launched three ways from the same state: kernel by kernel on
    one stream, kernel by kernel with each update beside the next tick's
    trace and sort (two tick records and active arrays alternating), and
    recorded once as a hipGraph replayed with each tick's record.  Runs
    differ only through the f32 summation order of the deposits, which moves
    fractions in the last bits and so, over ten ticks, which hit ends a few
    rays at the 1e-10 cut-off: full hit counts are equal, per-particle hits
    and last hits nearly everywhere, fractions and energies to rounding."""
    print_case("tick-graph", cfg, accum, walk)
    wl = make_workload(cfg, RATES, n=24 if cfg == 2 else 16, nray=256)
    ticks = [wl.rays(t) for t in range(1, 11)]
    pool = np.concatenate(ticks)
    firsts = np.concatenate([[0], np.cumsum([len(r) for r in ticks])])
    out = {}
    for way, graph, overlap in (("serial", 0, 0), ("overlap", 0, 1), ("graph", 1, 1)):
        with psr.Context(0) as c2:
            c2.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, accum)
            c2.set_tuning(psr.PSR_TUNE_DEPOSIT_WALK, walk)
            c2.set_tuning(psr.PSR_TUNE_TICK_GRAPH, graph)
            c2.set_tuning(psr.PSR_TUNE_TICK_OVERLAP, overlap)
            assert c2.tuning(psr.PSR_TUNE_TICK_OVERLAP) == overlap
            c2.set_tuning(psr.PSR_TUNE_TRACE_THREADS, 256 if cfg == 2 else 128)
            load(c2, wl)
            c2.rays_upload(pool)
            for t in range(1, 11):
                c2.tick(t, wl.tick_time, ray_first=int(firsts[t - 1]), nray=len(ticks[t - 1]))
            ps = wl.ps.copy()
            c2.state_download(ps)
            out[way] = (ps, c2.stats(), c2.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES))
    a, sa, ga = out["serial"]
    assert ga == 0 and out["overlap"][2] == 0 and out["graph"][2] >= 8  # the first tick(s) may size the hit slots
    for way in ("overlap", "graph"):
        b, sb, _ = out[way]
        assert sa["total_deposit_count"] == sb["total_deposit_count"], way
        # (measured: 0.1% of particles; a stale tick record would change most)
        assert np.mean(a.hits != b.hits) < 1e-2 and np.mean(a.lasthit != b.lasthit) < 1e-2, way
        for f in ("lambdaH", "lambdaHeI", "lambdaHeII"):
            va, vb = getattr(a, f), getattr(b, f)
            d = np.abs(va - vb)
            if wl.physics.flags.isothermal:
                # closed-form update: last-bit differences stay last-bit
                assert d.max() < 1e-5, (way, f)
            else:
                # an adaptive step decision can flip on a last-bit difference
                # (see test_full_tick_parity), which moves that particle's
                # result within the integrator's tolerance (rtol 1e-4);
                # bounded by the 1e-3 bar
                assert np.median(d) < 1e-9, (way, f)
                assert (d <= 1e-3 * np.maximum(np.abs(va), np.abs(vb)) + 1e-6).all(), (way, f)
        dr = np.abs(a.ie / b.ie - 1)
        assert np.median(dr) < 1e-6 and dr.max() < 1e-2, way


@pytest.mark.parametrize("cfg,n,nray", [(1, 32, 32), (2, 24, 256), (3, 16, 128)])
def test_full_tick_parity(ctx, cfg, n, nray):
    print_case("tick", cfg, n, nray)
    wl = make_workload(cfg, RATES, n=n, nray=nray)
    orc = oracle_for(wl)
    load(ctx, wl)
    out, ref, rays, _, off = run_one_tick(ctx, wl, orc, psr.PSR_PHASE_ALL)
    assert np.array_equal(out.lasthit, ref.lasthit)
    if wl.physics.flags.isothermal:
        # same algorithm, rates fixed per particle: agreement to rounding
        assert np.abs(out.lambdaH - ref.lambdaH).max() < 1e-6
        assert np.allclose(out.ie, ref.ie, rtol=1e-6)
    else:
        # log10(T) differs by an ulp between the device and libm, which can
        # flip an adaptive step decision: the bar is 1e-3 on fractions
        for f in ("lambdaH", "lambdaHeI", "lambdaHeII"):
            dx = np.abs(getattr(out, f) - getattr(ref, f))
            assert dx.max() < 1e-3 and np.median(dx) < 1e-9, f
        assert np.median(np.abs(out.ie / ref.ie - 1)) < 1e-9
        assert np.percentile(np.abs(out.ie / ref.ie - 1), 99) < 1e-2
    assert (out.yGdepHI == 0).all() and (out.heat == 0).all()


FLAG_CASES = {
    "h-caseA-iso": (1, dict(isothermal=True, on_the_spot=False, h_only=True)),
    "he-caseA-T": (3, dict(isothermal=False, on_the_spot=False, h_only=False)),
    "he-adiabatic": (3, dict(isothermal=False, adiabatic=True, h_only=False)),
    "he-comoving-z3": (3, dict(isothermal=False, h_only=False, comoving=True, redshift=3.0)),
    "h-fake-T": (1, dict(isothermal=True, h_only=True, fake_temperature=2.0e4)),
    "he-no-secondary": (3, dict(isothermal=False, h_only=False, secondary_ionization=False)),
}


@pytest.mark.parametrize("name", list(FLAG_CASES))
def test_full_tick_parity_flags(ctx, name):
    """the other run flags (case-A recombination with its counters, frozen
    energy, comoving scaling, a fixed temperature, no secondary ionization):
    a whole tick against the oracle"""
    print_case("tick flags", name)
    cfg, kw = FLAG_CASES[name]
    wl = make_workload(cfg, RATES, n=16, nray=64)
    u = wl.physics.units
    base = dict(isothermal=True, on_the_spot=True, h_only=True, secondary_ionization=True,
                comoving=False)
    base.update(kw)
    wl.physics = psr.Physics(RATES, u, psr.make_flags(**base), with_secion=not base["h_only"])
    orc = oracle_for(wl)
    load(ctx, wl)
    out, ref, rays, _, off = run_one_tick(ctx, wl, orc, psr.PSR_PHASE_ALL)
    assert np.array_equal(out.hits, ref.hits)
    assert np.array_equal(out.lasthit, ref.lasthit)
    fr = ("lambdaH",) if base["h_only"] else ("lambdaH", "lambdaHeI", "lambdaHeII")
    for f in fr + ("yGrecHII",):
        dx = np.abs(getattr(out, f).astype(np.float64) - getattr(ref, f).astype(np.float64))
        scale = 1.0 if f.startswith("lambda") else max(np.abs(getattr(ref, f)).max(), 1e-30)
        assert dx.max() / scale < 1e-3 and np.median(dx) / scale < 1e-6, f
    if base.get("adiabatic"):
        assert np.array_equal(out.ie, ref.ie)
    else:
        assert np.median(np.abs(out.ie / ref.ie - 1)) < 1e-6


def test_solve_steps_matches_oracle(ctx, golden):
    print_case("solve_steps")
    c = np.load(f"{golden}/chem_lsoda.npz")
    for key, hmf, iso in (("h_iso", 1.0, True), ("he", 0.76, False)):
        units = psr.default_units(hmf=hmf)
        flags = psr.make_flags(isothermal=iso, h_only=iso)
        phys = psr.Physics(RATES, units, flags)
        ctx.set_physics(phys)
        dev = ctx.solve_steps(c[f"{key}_in"])
        orc = po.Oracle(po.units(hmf=hmf), po.flags(isothermal=iso, h_only=iso),
                        kernel=po.reference_kernel_table())
        ref, bad = orc.step_evolve(c[f"{key}_in"])
        assert bad == 0
        tol = 1e-7 if iso else 1e-3
        for f in ("lambdaH", "lambdaHeI", "lambdaHeII"):
            dx = np.abs(dev[f] - ref[f])
            assert dx.max() < tol and np.median(dx) < 1e-9, f
        # against the reference's LSODA (fixture), the 1e-3 x_HII bar
        assert np.abs(dev["lambdaH"] - c[f"{key}_out"]["lambdaH"]).max() < 1e-3


@pytest.mark.slow
def test_iliev_multi_tick_ionized_fraction(ctx):
    """C1 geometry at 24^3 over 60 ticks: x_HII field within 1e-3 of the oracle"""
    print_case("iliev multi-tick")
    wl = make_workload(1, RATES, n=24, nray=32)
    orc = oracle_for(wl)
    load(ctx, wl)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    for tick in range(1, 61):
        rays = wl.rays(tick)
        ctx.rays_upload(rays)
        ctx.tick(tick, wl.tick_time)
        rc, _ = po.tick(orc, t, ref, rays, tick, wl.tick_time)
        assert rc == 0
    out = ctx.state_download(wl.ps.copy())
    dx = np.abs(out.lambdaH - ref.lambdaH)
    assert (1 - ref.lambdaH).max() > 0.5      # an ionized region formed
    assert dx.max() < 1e-3


def test_refused_ray_upload_leaves_no_pool(ctx):
    """a ray pool with a field index outside the tree is refused before any
    copy, and a tick over the refused pool returns PSR_ERR_ARG without a
    launch; a tree rebuilt with fewer fields refuses the older pool too"""
    print_case("refused ray upload")
    wl = make_workload(5, RATES, n=None, nray=None, nfield=3)
    load(ctx, wl)
    rays = wl.rays(1)
    assert int(rays["field"].max()) == 2
    bad = rays.copy()
    bad["field"][len(bad) // 2] = 3
    with pytest.raises(psr.PsrError) as e:
        ctx.rays_upload(bad)
    assert e.value.code == 1
    rc = ctx._lib.psr_tick(ctx.h, 0, len(rays), 1, wl.tick_time, psr.PSR_PHASE_ALL, psr.PSR_CHEM_EXACT)
    assert rc == 1
    ctx.rays_upload(rays)
    ctx.tick(1, wl.tick_time, nray=len(rays))
    ctx.sync()
    # the same pool after a rebuild with one field only
    one = make_workload(1, RATES, n=8, nray=8)
    load(ctx, one)
    rc = ctx._lib.psr_tick(ctx.h, 0, len(rays), 2, wl.tick_time, psr.PSR_PHASE_ALL, psr.PSR_CHEM_EXACT)
    assert rc == 1


def test_void_rate_table_steps_on_the_device(ctx):
    """psr_solve_steps with the reference's void rate table: the fractions
    lose exactly their deposits (x = x0 - yGdep, He lines included) and a
    neutral particle gains exactly its heat, as the oracle pins show"""
    import os
    from conftest import GOLDEN
    print_case("void rate table solve_steps")
    vp = os.path.join(GOLDEN, "atomic_rates_void.txt")
    for ots in (True, False):
        u = psr.default_units(hmf=0.76)
        ctx.set_physics(psr.Physics(vp, u, psr.make_flags(isothermal=False, h_only=False,
                                                          on_the_spot=ots)))
        phys = psr.Physics(vp, u)
        s = np.zeros(2, psr.STEP_DTYPE)
        nH = 1e-3 / u.U_CM ** 3
        s["nH"] = nH
        s["T"] = 1e4
        s["time"] = 3.0
        for k, v in dict(lambdaH=0.7, lambdaHeI=0.5, lambdaHeII=0.3, yGdepHI=0.2, yGdepHeI=0.1,
                         yGdepHeII=0.05).items():
            s[k][0] = v
        s["ie"][0] = phys.Tye2ie(1e4, 0.5)
        ie0 = phys.Tye2ie(1e4, 0.0)
        s["lambdaH"][1], s["lambdaHeI"][1], s["lambdaHeII"][1] = 1.0, 1.0, 0.0
        s["ie"][1] = ie0
        s["heat"][1] = 0.37 * ie0
        o = ctx.solve_steps(s)
        assert abs(o["lambdaH"][0] - 0.5) < 1e-14
        assert abs(o["lambdaHeI"][0] - 0.4) < 1e-14
        assert abs(o["lambdaHeII"][0] - 0.25) < 1e-14
        assert o["ie"][1] == pytest.approx(ie0 * 1.37, rel=1e-14)
        assert o["lambdaH"][1] == 1.0 and o["lambdaHeI"][1] == 1.0


def _tick_until_decided(c, wl, rays, limit=40):
    c.rays_upload(rays)
    for t in range(1, limit + 1):
        c.tick(t, wl.tick_time)
        c.sync()
        if (c.tuning(psr.PSR_TUNE_DEPOSIT_WALK) != psr.PSR_WALK_AUTO
                and c.tuning(psr.PSR_TUNE_TICK_GRAPH) in (0, 1)
                and c.tuning(psr.PSR_TUNE_TRACE_THREADS) in (128, 256)):
            return t
    raise AssertionError("the autotuner did not decide")


def test_tuning_persistence(tmp_path, monkeypatch):
    """decided launch choices saved by one context and loaded by another of
    the same workload (no trial ticks: the first tick already runs the
    choice); another workload finds no entry; a tree rebuild restarts the
    tuning; PSR_TUNING_FILE does the same across contexts by itself"""
    print_case("tuning persistence")
    wl = make_workload(2, RATES, n=24, nray=256)
    rays = wl.rays(1)
    f = tmp_path / "tuning.txt"
    with psr.Context(0) as a:
        load(a, wl)
        _tick_until_decided(a, wl, rays)
        choice = [a.tuning(p) for p in (psr.PSR_TUNE_DEPOSIT_ACCUM, psr.PSR_TUNE_DEPOSIT_WALK,
                                        psr.PSR_TUNE_TICK_GRAPH, psr.PSR_TUNE_TRACE_THREADS)]
        a.tuning_save(f)
        a.tuning_save(f)                       # the workload's line is replaced, not repeated
    lines = f.read_text().splitlines()
    assert len(lines) == 1 and "npar=%d" % wl.ps.npar in lines[0] and "dep=" in lines[0]
    with psr.Context(0) as b:
        load(b, wl)
        b.tuning_load(f)
        assert [b.tuning(p) for p in (psr.PSR_TUNE_DEPOSIT_ACCUM, psr.PSR_TUNE_DEPOSIT_WALK,
                                      psr.PSR_TUNE_TICK_GRAPH, psr.PSR_TUNE_TRACE_THREADS)] == choice
        b.rays_upload(rays)
        b.tick(1, wl.tick_time)
        b.sync()
        if choice[2] == 1:
            assert b.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES) == 1
        load(b, wl)                            # a rebuilt tree restarts the trials
        assert b.tuning(psr.PSR_TUNE_DEPOSIT_WALK) == psr.PSR_WALK_AUTO
    other = make_workload(2, RATES, n=16, nray=64)
    with psr.Context(0) as c:
        load(c, other)
        with pytest.raises(psr.PsrError):
            c.tuning_load(f)
    env = tmp_path / "env.txt"
    monkeypatch.setenv("PSR_TUNING_FILE", str(env))
    with psr.Context(0) as d:
        load(d, wl)
        _tick_until_decided(d, wl, rays)
        want = d.tuning(psr.PSR_TUNE_DEPOSIT_WALK)
    assert env.exists() and len(env.read_text().splitlines()) == 1
    with psr.Context(0) as e:
        load(e, wl)
        e.rays_upload(rays)
        e.tick(1, wl.tick_time)
        e.sync()
        assert e.tuning(psr.PSR_TUNE_DEPOSIT_WALK) == want
