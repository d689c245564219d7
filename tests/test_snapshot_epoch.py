"""psphray snapshots and epochs (src/reader-psphray.inc; src/psystem.c:221-504,
:599-673): the library's host code (psr_snapshot_*, psr_epoch_load,
psr_epoch_switch) against the oracle's numpy/Python restatement (oracle/epoch.py)
bit for bit, against hand-counted file sizes and against invariants of the
ID match; and, on the GPU, ticks that continue across an epoch switch
against the oracle doing the same."""
import copy

import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd import _abi
from psphray_amd import snapshot as snap
from oracle import epoch as oe

from conftest import RATES, print_case

FIELDS = ("pos", "flag", "lasthit", "hits", "sml", "mass", "rho", "ie", "yeMET", "yGdepHI", "yGdepHeI",
          "yGdepHeII", "yGrecHII", "yGrecHeII", "yGrecHeIII", "heat", "lambdaH", "lambdaHeI", "lambdaHeII")


def test_header_layout_matches_format():
    """psr_snap_header is the format's 256-byte record: every field at the
    offset natural C alignment gives struct header_t"""
    assert _abi.C.sizeof(_abi.SnapHeader) == 256
    for name in oe.HEADER_DTYPE.names:
        assert getattr(_abi.SnapHeader, name).offset == oe.HEADER_DTYPE.fields[name][1], name


def _blocks(rng, ngas, nsrc):
    n = ngas + nsrc
    return {"pos": rng.random((n, 3), np.float32), "id": rng.integers(0, 2**40, n).astype(np.uint64),
            "mass": rng.random(n, np.float32), "ie": rng.random(ngas, np.float32),
            "rho": rng.random(ngas, np.float32), "ye": rng.random(ngas, np.float32),
            "xHI": rng.random(ngas, np.float32), "sml": rng.random(ngas, np.float32),
            "xHeI": rng.random(ngas, np.float32), "gammaHI": rng.random(ngas, np.float32),
            "lasthit": rng.integers(-5, 100, ngas), "hits": rng.integers(0, 9, ngas).astype(np.int32),
            "ngammas": rng.random(nsrc)}


def _oracle_header(hd):
    return np.frombuffer(bytes(hd), oe.HEADER_DTYPE)[0]


def test_snapshot_sizes_by_hand(tmp_path):
    """5 gas + 2 sources, helium and gammaHI on: header 264; pos 7*12, id
    7*8, mass 7*4; ie, rho, ye, xHI, sml 5*4 each; xHeI, xHeII 5*4;
    gammaHI 5*4; lasthit 5*8; hits 5*4; ngammas, spec 2*8; plus 8 bytes of
    framing per block: 804 bytes"""
    hd = snap.make_header(5, 2, boxsize=1.0)
    blocks = _blocks(np.random.default_rng(0), 5, 2)
    snap.write(tmp_path / "a", hd, blocks)
    oe.write_snapshot(tmp_path / "b", _oracle_header(hd), blocks)
    assert (tmp_path / "a").stat().st_size == (tmp_path / "b").stat().st_size == 804
    assert (tmp_path / "a").read_bytes() == (tmp_path / "b").read_bytes()
    # offsets from the library agree with the framing
    assert snap.block_info(hd, "pos") == (7, 12, 264)
    assert snap.block_info(hd, "hits")[2] == 804 - 24 - 24 - (20 + 8)
    assert snap.block_info(hd, "Hmf")[0] == 0 and snap.block_info(hd, "sft")[0] == 0


@pytest.mark.parametrize("ngas,nsrc,flags", [(0, 0, {}), (1, 0, {}), (37, 3, dict(flag_Hmf=1, flag_eos=1)),
                                             (1000, 0, dict(flag_Helium=0, flag_gammaHI=0)),
                                             (64, 5, dict(flag_cloudy=1, flag_Hemf=1))])
def test_snapshot_round_trips(tmp_path, ngas, nsrc, flags):
    """library writer -> oracle reader and oracle writer -> library reader
    give back every block; the two writers write the same bytes"""
    rng = np.random.default_rng(ngas)
    hd = snap.make_header(ngas, nsrc, boxsize=3.5, redshift=6.0)
    for k, v in flags.items():
        setattr(hd, k, v)
    oh = _oracle_header(hd)
    blocks = {}
    for name, item, dt, k, types, flag in oe.BLOCKS:
        ln = oe.block_length(oh, types, flag)
        if ln:
            a = rng.integers(0, 255, ln * item, dtype=np.uint8).view(dt)
            blocks[name] = a.reshape(ln, 3) if k == 3 else a
    snap.write(tmp_path / "lib", hd, blocks)
    oe.write_snapshot(tmp_path / "orc", oh, blocks)
    assert (tmp_path / "lib").read_bytes() == (tmp_path / "orc").read_bytes()
    h1, got = snap.read(tmp_path / "orc")
    assert bytes(h1) == bytes(hd)
    assert set(got) == set(blocks)
    for name in blocks:
        assert got[name].tobytes() == np.ascontiguousarray(blocks[name]).tobytes(), name
    h2, got2 = oe.read_snapshot(tmp_path / "lib")
    for name in blocks:
        assert got2[name].tobytes() == np.ascontiguousarray(blocks[name]).tobytes(), name


@pytest.mark.parametrize("fmt", ["massiveblack", "e5"])
@pytest.mark.parametrize("N,flags", [((30, 50, 0, 0, 7, 2), {}), ((30, 0, 0, 0, 0, 0), dict(flag_met=0)),
                                     ((5, 9, 0, 0, 4, 3), dict(flag_sfr=0, flag_sft=0))])
def test_massiveblack_round_trips(tmp_path, fmt, N, flags):
    """the massiveblack table (dark matter in pos/vel/id only, metals of gas
    and stars, black holes' mass and accretion): library and oracle write
    the same bytes and read each other's files; gas_blocks takes type 0's
    rows"""
    rng = np.random.default_rng(sum(N))
    hd = snap.make_header(N[0])
    for t in range(6):
        hd.N[t] = N[t]
    for k, v in flags.items():
        setattr(hd, k, v)
    oh = _oracle_header(hd)
    blocks = {}
    for name, item, dt, k, types, flag in oe.BLOCKS_MASSIVEBLACK:
        ln = oe.block_length(oh, types, flag)
        if ln:
            a = rng.integers(0, 255, ln * item, dtype=np.uint8).view(dt)
            blocks[name] = a.reshape(ln, 3) if k == 3 else a
    assert ("sfr" in blocks) == bool(hd.flag_sfr) and "vel" in blocks
    snap.write(tmp_path / "lib", hd, blocks, fmt=fmt)
    oe.write_snapshot(tmp_path / "orc", oh, blocks, fmt="massiveblack")
    assert (tmp_path / "lib").read_bytes() == (tmp_path / "orc").read_bytes()
    _, got = snap.read(tmp_path / "orc", fmt=fmt)
    assert set(got) == set(blocks)
    for name in blocks:
        assert got[name].tobytes() == np.ascontiguousarray(blocks[name]).tobytes(), name
    _, gas = snap.gas_blocks(tmp_path / "lib", fmt=fmt)
    assert len(gas["id"]) == N[0] and gas["pos"].tobytes() == blocks["pos"][:N[0]].tobytes()
    with pytest.raises(ValueError):
        snap.read(tmp_path / "lib", fmt="gadget")
    with pytest.raises(psr.PsrError):
        snap.block_info(hd, "lasthit", fmt)           # a psphray block


def test_snapshot_missing_blocks_are_zero_and_bad_files_refused(tmp_path):
    hd = snap.make_header(4, 0)
    snap.write(tmp_path / "a", hd, {"pos": np.ones((4, 3))})
    _, got = snap.read(tmp_path / "a")
    assert (got["pos"] == 1).all() and (got["rho"] == 0).all() and (got["id"] == 0).all()
    with pytest.raises(ValueError):
        snap.write(tmp_path / "b", hd, {"pos": np.ones((5, 3))})
    (tmp_path / "c").write_bytes(b"\0" * 100)
    with pytest.raises(psr.PsrError):
        snap.read(tmp_path / "c")
    raw = bytearray((tmp_path / "a").read_bytes())
    raw[264] ^= 1                                      # the pos record's size marker
    (tmp_path / "d").write_bytes(bytes(raw))
    with pytest.raises(psr.PsrError):
        snap.read(tmp_path / "d", names=("pos",))


def test_write_output_blocks(tmp_path):
    """write_output writes what psystem_write_output writes: x_HI, x_HeI,
    x_HeII from the lambdas, ye of H and He only, yGrecHII as gammaHI, the
    sources' positions as type-5 particles and their Ngamma_dot"""
    rng = np.random.default_rng(5)
    u = psr.default_units()
    n = 50
    ps = psr.ParticleSystem(n)
    ps.pos[:] = rng.random((n, 3))
    for f in ("sml", "mass", "rho", "ie", "yeMET", "yGrecHII"):
        getattr(ps, f)[:] = rng.random(n)
    ps.lambdaH[:] = rng.random(n)
    ps.lambdaHeI[:] = rng.random(n) * 0.5
    ps.lambdaHeII[:] = rng.random(n) * 0.5
    ps.lasthit[:] = rng.integers(0, 10, n)
    ps.hits[:] = rng.integers(0, 10, n)
    ids = np.arange(n, dtype=np.uint64) * 7
    spos = rng.random((2, 3))
    sng = rng.random(2)
    snap.write_output(tmp_path / "o", ps, ids, u, 2.0, redshift=3.0, src_pos=spos, src_ngd=sng)
    h, b = oe.read_snapshot(tmp_path / "o")
    assert tuple(h["N"]) == (n, 0, 0, 0, 0, 2) and h["flag_Helium"] == 1 and h["flag_gammaHI"] == 1
    assert h["boxsize"] == 2.0 and h["redshift"] == 3.0
    assert np.array_equal(b["pos"][:n], ps.pos) and np.array_equal(b["pos"][n:], spos.astype(np.float32))
    assert np.array_equal(b["id"][:n], ids)
    assert np.array_equal(b["xHI"], ps.lambdaH.astype(np.float32))
    assert np.array_equal(b["xHeII"], ps.lambdaHeII.astype(np.float32))
    xHeIII = 1.0 - ps.lambdaHeI - ps.lambdaHeII
    ye = (1 - ps.lambdaH) + u.C_HEMF / u.C_HMF * 0.25 * (ps.lambdaHeII + 2 * xHeIII)
    assert np.array_equal(b["ye"], ye.astype(np.float32))
    assert np.array_equal(b["gammaHI"], ps.yGrecHII) and np.array_equal(b["ngammas"], sng)
    assert np.array_equal(b["lasthit"], ps.lasthit) and np.array_equal(b["hits"], ps.hits)


# ------------------------------------------------------------------ epochs
def _gas(rng, n, box=100.0, ids=None):
    g = {"id": np.arange(n, dtype=np.uint64) if ids is None else ids,
         "pos": (rng.random((n, 3)) * box).astype(np.float32),
         "mass": rng.random(n).astype(np.float32) + 0.5, "sml": rng.random(n).astype(np.float32) + 1,
         "rho": rng.random(n).astype(np.float32), "ie": rng.random(n).astype(np.float32),
         # x_HI and ye spanning every clamp of the ionization set-up
         "xHI": rng.choice([-0.2, 0.0, 0.3, 0.999, 1.0, 1.3], n).astype(np.float32),
         "ye": rng.choice([0.0, 0.1, 0.7, 1.05, 1.2, 1.5, 3.0], n).astype(np.float32)}
    return g


def _oracle_psys(ps, ids):
    o = oe.Psys(len(ps.mass))
    o.npar = ps.npar
    for f in FIELDS:
        getattr(o, f)[:] = getattr(ps, f)
    o.id[:] = ids
    return o


def _same(ps, ids, o, what):
    assert ps.npar == o.npar, what
    for f in FIELDS:
        a, b = getattr(ps, f), getattr(o, f)
        assert a.tobytes() == b.tobytes(), (what, f, np.nonzero(a != b)[0][:5])
    assert np.array_equal(ids, o.id), what


@pytest.mark.parametrize("h_only", [0, 1])
def test_epoch_load_equals_oracle(h_only):
    rng = np.random.default_rng(10 + h_only)
    u = psr.default_units(hmf=0.76)
    g = _gas(rng, 400)
    ps, ids = snap.epoch_load(g, u, h_only)
    assert len(ps.mass) == int(400 * 1.05) and ps.npar == 400
    o = oe.Psys(len(ps.mass))
    oe.read_epoch(o, g, u.C_HMF, h_only)
    _same(ps, ids, o, "load")
    # x_HI outside [0, 1] is pinned; every fraction is a fraction
    assert ((ps.lambdaH >= 0) & (ps.lambdaH <= 1)).all()
    if not h_only:
        assert ((ps.lambdaHeI >= 0) & (ps.lambdaHeI <= 1)).all()
    with pytest.raises(psr.PsrError):
        snap.epoch_load(g, u, h_only, npar_max=399)


def _next_snapshot(rng, g, n_new, drop, jitter, collide, box=100.0):
    """the next epoch's gas: most particles moved a little, some dropped, new
    IDs added, some IDs equal to an old ID plus a multiple of 2^28 (same hash
    bucket, a different particle)"""
    n = len(g["id"])
    keep = np.sort(rng.choice(n, n - drop, replace=False))
    ids = g["id"][keep]
    pos = g["pos"][keep] + (rng.standard_normal((len(keep), 3)) * jitter).astype(np.float32)
    new_ids = np.arange(n_new, dtype=np.uint64) + np.uint64(10 ** 6)
    col = rng.choice(n, collide, replace=False)
    new_ids[:collide] = g["id"][col] + np.uint64(1 << 28) * rng.integers(1, 4, collide).astype(np.uint64)
    new_pos = (rng.random((n_new, 3)) * box).astype(np.float32)
    new_pos[:collide // 2] = g["pos"][col[:collide // 2]]   # a collision sitting on the old particle
    out = _gas(rng, len(keep) + n_new, box=box, ids=np.concatenate([ids, new_ids]))
    out["pos"] = np.concatenate([pos, new_pos])
    perm = rng.permutation(len(out["id"]))
    return {k: v[perm] for k, v in out.items()}


@pytest.mark.parametrize("h_only,periodic,collide", [(0, 0, 20), (1, 0, 20), (0, 1, 20), (0, 0, 0), (1, 1, 0)])
def test_epoch_switch_equals_oracle(h_only, periodic, collide):
    """the library's match (bucket sort) against the oracle's (the
    reference's hash chains): every byte of every array up to npar_max, the
    stale rows past npar included, and the statistics.  Without hash
    collisions every snapshot particle is present once; an ID in an old
    particle's bucket within 500 of it takes that particle's slot (the
    reference's match), so collisions merge particles."""
    rng = np.random.default_rng(20 + h_only + 2 * periodic + collide)
    u = psr.default_units(hmf=0.76)
    g = _gas(rng, 600)
    # old particles sharing a bucket at one position: the first visited of
    # equally near candidates wins (the hash chain's order, last inserted first)
    if collide:
        g["id"][580:] = g["id"][:20] + np.uint64(1 << 28)
        g["pos"][580:] = g["pos"][:20]
    ps, ids = snap.epoch_load(g, u, h_only, npar_max=700)
    # state evolved by an epoch of ticks
    for f in ("yGdepHI", "yGrecHII", "heat"):
        getattr(ps, f)[:ps.npar] = rng.random(ps.npar)
    ps.lasthit[:ps.npar] = rng.integers(0, 50, ps.npar)
    ps.hits[:ps.npar] = rng.integers(0, 9, ps.npar)
    nxt = _next_snapshot(rng, g, n_new=60, drop=90, jitter=2.0, collide=collide)
    if periodic:
        nxt["pos"] = np.mod(nxt["pos"], np.float32(100.0)).astype(np.float32)
    o = _oracle_psys(ps, ids)
    st = snap.epoch_switch(ps, ids, nxt, u, h_only, prev_duration=5.0, prev_nticks=50, tick_time=0.3,
                           periodic=periodic, boxsize=100.0)
    ost = oe.switch_epoch(o, nxt, u.C_HMF, h_only, 5.0, 50, 0.3, periodic, 100.0)
    _same(ps, ids, o, "switch")
    assert st["added"] == ost["added"] and st["lost"] == 0 and st["removed"] == ost["removed"]
    assert st["matched"] == len(nxt["id"]) - ost["added"]
    assert st["mean_shift"] == pytest.approx(ost["distsum"] / len(nxt["id"]), rel=1e-12)
    assert (ps.flag[:ps.npar] == 0).all()
    # a matched slot keeps its old ID (the match compares hashes)
    mask = np.uint64(oe.IDHASHMASK)
    assert set(ids[:ps.npar] & mask) <= set(nxt["id"] & mask)
    if collide:
        assert ps.npar < len(nxt["id"])
        return
    # every snapshot particle is present exactly once, at its snapshot position
    assert ps.npar == len(nxt["id"])
    order = np.argsort(ids[:ps.npar], kind="stable")
    so = np.argsort(nxt["id"], kind="stable")
    assert np.array_equal(ids[:ps.npar][order], nxt["id"][so])
    assert set(map(bytes, ps.pos[:ps.npar])) == set(map(bytes, nxt["pos"]))
    assert (ps.flag[:ps.npar] == 0).all()


def test_epoch_switch_unchanged_snapshot_is_identity():
    """switching to the same snapshot matches every particle to itself:
    only lasthit changes, to (lasthit * dt_old - duration) / dt_new"""
    rng = np.random.default_rng(3)
    u = psr.default_units(hmf=0.76)
    g = _gas(rng, 300)
    ps, ids = snap.epoch_load(g, u, 0)
    ps.lasthit[:300] = rng.integers(0, 40, 300)
    ps.yGrecHII[:300] = rng.random(300)
    before = ps.copy()
    st = snap.epoch_switch(ps, ids, g, u, 0, prev_duration=4.0, prev_nticks=40, tick_time=0.05)
    assert st == dict(matched=300, added=0, lost=0, removed=0, mean_shift=0.0)
    for f in FIELDS:
        if f != "lasthit":
            assert np.array_equal(getattr(ps, f), getattr(before, f)), f
    assert np.array_equal(ps.lasthit[:300], ((before.lasthit[:300] * 0.1 - 4.0) / 0.05).astype(np.int64))


def test_epoch_switch_far_particle_is_new_and_capacity():
    """an ID whose old particle is 500 or more away is a new particle with
    the snapshot's state (the old one is removed); unmatched particles beyond
    npar_max are reported"""
    u = psr.default_units(hmf=1.0)
    g = {"id": np.array([1, 2], np.uint64), "pos": np.array([[0, 0, 0], [10, 0, 0]], np.float32),
         "mass": np.ones(2, np.float32), "sml": np.ones(2, np.float32), "rho": np.ones(2, np.float32),
         "ie": np.ones(2, np.float32), "xHI": np.ones(2, np.float32), "ye": np.zeros(2, np.float32)}
    ps, ids = snap.epoch_load(g, u, 1, npar_max=8)
    ps.lambdaH[:2] = 0.25
    nxt = dict(g, pos=np.array([[499.9, 0, 0], [510.0, 0, 0]], np.float32), xHI=np.full(2, 0.5, np.float32))
    st = snap.epoch_switch(ps, ids, nxt, u, 1, 1.0, 1, 1.0)
    assert (st["matched"], st["added"], st["removed"]) == (1, 1, 1)
    assert ps.npar == 2
    assert sorted(ps.lambdaH[:2]) == [0.25, 0.5]
    small = {k: v for k, v in g.items()}
    ps2, ids2 = snap.epoch_load(small, u, 1, npar_max=2)
    many = _gas(np.random.default_rng(1), 5, ids=np.arange(10, 15, dtype=np.uint64))
    with pytest.raises(psr.PsrError):
        snap.epoch_switch(ps2, ids2, many, u, 1, 1.0, 1, 1.0)


# ------------------------------------------------------------ on the device
def _view(ps, ids=None):
    """a ParticleSystem over the first ps.npar rows (no copy)"""
    v = psr.ParticleSystem(0)
    for f in ps.fields():
        setattr(v, f, getattr(ps, f)[: ps.npar])
    v.npar = ps.npar
    return v


@pytest.mark.gpu
def test_ticks_across_epoch_switch():
    """two ticks, an epoch switch (particles moved, removed and added; the
    tree rebuilt on the device; lasthit in the new ticks), two more ticks:
    the device against the oracle doing the same, hits and lasthit exact"""
    print_case("ticks across an epoch switch")
    from psphray_amd.fields import make_workload
    from oracle import pyoracle as po
    from test_gpu_parity import oracle_for
    wl = make_workload(2, RATES, n=16, nray=512)
    em2 = copy.deepcopy(wl.emitter)
    n0 = wl.ps.npar
    cap = int(n0 * 1.1)
    ps = snap.resized(wl.ps, cap)
    ids = np.zeros(cap, np.uint64)
    ids[:n0] = np.arange(n0, dtype=np.uint64) * 3 + 1
    ref = ps.copy()
    orc = oracle_for(wl)
    rng = np.random.default_rng(7)
    u = wl.physics.units
    with psr.Context(0) as ctx:
        ctx.set_physics(wl.physics)
        ctx.load(_view(ps), boxsize=wl.boxsize)
        t = po.Tree(ref.pos[:n0], ref.sml[:n0], wl.boxsize)
        for tk in (1, 2):
            rays = wl.rays(tk)
            ctx.rays_upload(rays)
            ctx.tick(tk, wl.tick_time, nray=len(rays))
            rc, _ = po.tick(orc, t, _view(ref), rays, tk, wl.tick_time)
            assert rc == 0
        ctx.state_download(_view(ps))
        assert np.array_equal(ps.hits, ref.hits) and np.array_equal(ps.lasthit, ref.lasthit)
        # the next epoch's snapshot
        g = {"id": ids[:n0].copy(), "pos": ps.pos[:n0].copy(), "mass": ps.mass[:n0].copy(),
             "sml": ps.sml[:n0].copy(), "rho": ps.rho[:n0].copy(), "ie": ps.ie[:n0].copy(),
             "xHI": np.ones(n0, np.float32), "ye": np.zeros(n0, np.float32)}
        nxt = _next_snapshot(rng, g, n_new=n0 // 20, drop=n0 // 10, jitter=1e-3 * wl.boxsize, collide=8,
                             box=0.999 * wl.boxsize)
        nxt["pos"] = np.clip(nxt["pos"], 0, np.float32(wl.boxsize * 0.999)).astype(np.float32)
        nxt["sml"][:] = np.median(ps.sml[:n0])
        nxt["mass"][:] = ps.mass[0]
        ref_ids = ids.copy()
        st = snap.epoch_switch(ps, ids, nxt, u, 1, prev_duration=2 * wl.tick_time, prev_nticks=2,
                               tick_time=wl.tick_time)
        rst = snap.epoch_switch(ref, ref_ids, nxt, u, 1, prev_duration=2 * wl.tick_time, prev_nticks=2,
                                tick_time=wl.tick_time)
        assert st == rst and st["added"] > 0 and st["removed"] > 0
        assert np.array_equal(ids, ref_ids) and ps.npar == ref.npar
        n1 = ps.npar
        ctx.load(_view(ps), boxsize=wl.boxsize)
        t = po.Tree(ref.pos[:n1], ref.sml[:n1], wl.boxsize)
        for tk in (1, 2):
            rays = em2.emit(tk, wl.nray)
            ctx.rays_upload(rays)
            ctx.tick(tk, wl.tick_time, nray=len(rays))
            rc, _ = po.tick(orc, t, _view(ref), rays, tk, wl.tick_time)
            assert rc == 0
        ctx.state_download(_view(ps))
    assert np.array_equal(ps.hits[:n1], ref.hits[:n1])
    assert np.array_equal(ps.lasthit[:n1], ref.lasthit[:n1])
    assert (ps.lasthit[:n1] <= 2).all() and (ps.hits[:n1] > 0).any()
    assert np.abs(ps.lambdaH[:n1] - ref.lambdaH[:n1]).max() < 1e-6


@pytest.mark.gpu
def test_output_gather_and_device_writer(tmp_path):
    """psr_output_gather after two H + He ticks: each block is the state's
    value as psystem_write_output derives it (fractions and ye as floats,
    yGrecHII as gammaHI, ie, lasthit, hits), and a snapshot written from
    it is byte for byte the one written from the downloaded state"""
    print_case("output gather")
    from psphray_amd.fields import make_workload
    wl = make_workload(3, RATES, n=16, nray=256)
    u = wl.physics.units
    ids = np.arange(wl.ps.npar, dtype=np.uint64) * 5 + 2
    with psr.Context(0) as ctx:
        ctx.set_physics(wl.physics)
        ctx.load(wl.ps, boxsize=wl.boxsize)
        for t in (1, 2):
            rays = wl.rays(t)
            ctx.rays_upload(rays)
            ctx.tick(t, wl.tick_time, nray=len(rays))
        g = ctx.output_gather(wl.ps.npar)
        st = ctx.state_download(wl.ps.copy())
        snap.write_output_device(tmp_path / "dev", ctx, wl.ps, ids, u, wl.boxsize, redshift=6.0,
                                 src_pos=wl.emitter.pos, src_ngd=wl.emitter.ngd)
    assert (st.hits > 0).any()
    assert np.array_equal(g["xHI"], st.lambdaH.astype(np.float32))
    assert np.array_equal(g["xHeI"], st.lambdaHeI.astype(np.float32))
    assert np.array_equal(g["xHeII"], st.lambdaHeII.astype(np.float32))
    assert np.array_equal(g["ye"], snap.ye_of(st, u).astype(np.float32))
    assert np.array_equal(g["gammaHI"], st.yGrecHII) and np.array_equal(g["ie"], st.ie)
    assert np.array_equal(g["lasthit"], st.lasthit) and np.array_equal(g["hits"], st.hits)
    snap.write_output(tmp_path / "host", st, ids, u, wl.boxsize, redshift=6.0, src_pos=wl.emitter.pos,
                      src_ngd=wl.emitter.ngd)
    assert (tmp_path / "dev").read_bytes() == (tmp_path / "host").read_bytes()
