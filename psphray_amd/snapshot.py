"""Snapshots and epoch switching (host side).

read / write: the psphray snapshot format (src/reader-psphray.inc) and the
massiveblack one (src/reader-massiveblack.inc) through psr_snapshot_*: a
256-byte header record, then the blocks in the format's order, each framed
by its byte count.  write_output mirrors the reference's
psystem_write_output (src/psystem.c:599-673): gas positions, sml, rho, mass,
x_HI, x_HeI, x_HeII, ye, yGrecHII (as gammaHI), ie, lasthit, hits, ids, then
the sources as type-5 particles (positions, Ngamma_dot).

epoch_switch: psystem_match_epoch + compress (src/psystem.c:381-504,
:109-150) through psr_epoch_switch, on a ParticleSystem with spare capacity.
"""
import ctypes as C

import numpy as np

from . import ParticleSystem, PsrError, _abi

BLOCK_DTYPE = {
    "pos": ("<f4", 3), "id": ("<u8", 1), "mass": ("<f4", 1), "ie": ("<f4", 1), "rho": ("<f4", 1),
    "ye": ("<f4", 1), "xHI": ("<f4", 1), "sml": ("<f4", 1), "Hmf": ("<f4", 1), "Hemf": ("<f4", 1),
    "xHeI": ("<f4", 1), "xHeII": ("<f4", 1), "gammaHI": ("<f4", 1), "HIa_cloudy": ("<f4", 1),
    "eos": ("<f4", 1), "lasthit": ("<i8", 1), "hits": ("<i4", 1), "sft": ("<f4", 1),
    "ngammas": ("<f8", 1), "spec": ("S8", 1),
}
BLOCK_DTYPE_MASSIVEBLACK = {
    "pos": ("<f4", 3), "vel": ("<f4", 3), "id": ("<u8", 1), "mass": ("<f4", 1), "ie": ("<f4", 1),
    "rho": ("<f4", 1), "ye": ("<f4", 1), "xHI": ("<f4", 1), "sml": ("<f4", 1), "sfr": ("<f4", 1),
    "sft": ("<f4", 1), "met": ("<f4", 1), "bhmass": ("<f4", 1), "bhmdot": ("<f4", 1),
}
FORMATS = {"psphray": BLOCK_DTYPE, "massiveblack": BLOCK_DTYPE_MASSIVEBLACK, "e5": BLOCK_DTYPE_MASSIVEBLACK}


def _dtypes(fmt):
    if fmt not in FORMATS:
        raise ValueError(f"unknown snapshot format {fmt}")
    return FORMATS[fmt]


def _err():
    return C.create_string_buffer(512)


def block_info(hdr, name, fmt="psphray"):
    """(length, item size, offset) of a block under a header"""
    ln, it, off = C.c_int64(), C.c_int64(), C.c_int64()
    rc = _abi.lib().psr_snapshot_block_info(fmt.encode(), C.byref(hdr), name.encode(), C.byref(ln), C.byref(it),
                                            C.byref(off))
    if rc:
        raise PsrError(rc, f"unknown block {name}")
    return ln.value, it.value, off.value


def make_header(ngas, nsrc=0, boxsize=0.0, redshift=0.0, time=1.0, h=0.72, helium=True, gammaHI=True,
                nfiles=1):
    """a header for ngas gas and nsrc source particles (the flags the
    reference's writer sets, src/psystem.c:621-636)"""
    hd = _abi.SnapHeader()
    hd.N[0] = ngas
    hd.N[5] = nsrc
    hd.Nparticle_total_low[0] = ngas & 0xFFFFFFFF
    hd.Nparticle_total_high[0] = ngas >> 32
    hd.Nparticle_total_low[5] = nsrc & 0xFFFFFFFF
    hd.Nparticle_total_high[5] = nsrc >> 32
    hd.boxsize = boxsize
    hd.redshift = redshift
    hd.time = time
    hd.h = h
    hd.Nfiles = nfiles
    hd.flag_Helium = int(helium)
    hd.flag_gammaHI = int(gammaHI)
    hd.flag_cool = hd.flag_met = hd.flag_sft = hd.flag_sfr = hd.flag_feedback = 1
    return hd


def write(path, hdr, blocks, fmt="psphray"):
    """header + blocks (name -> array); blocks the header calls for and
    not given are written as zeros"""
    names, bufs, keep = [], [], []
    for name, arr in blocks.items():
        dt, k = _dtypes(fmt)[name]
        ln, _, _ = block_info(hdr, name, fmt)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.size != ln * k:
            raise ValueError(f"block {name}: {a.size} values, the header asks for {ln * k}")
        keep.append(a)
        names.append(name.encode())
        bufs.append(a.ctypes.data)
    n = len(names)
    cn = (C.c_char_p * max(n, 1))(*names)
    cb = (C.c_void_p * max(n, 1))(*bufs)
    err = _err()
    rc = _abi.lib().psr_snapshot_write(fmt.encode(), str(path).encode(), C.byref(hdr), cn, cb, n, err, 512)
    if rc:
        raise PsrError(rc, err.value.decode())


def read(path, names=None, fmt="psphray"):
    """(header, {block: array}) for every present block (or those named)"""
    hdr = _abi.SnapHeader()
    err = _err()
    rc = _abi.lib().psr_snapshot_read_header(str(path).encode(), C.byref(hdr), err, 512)
    if rc:
        raise PsrError(rc, err.value.decode())
    out = {}
    for name, (dt, k) in _dtypes(fmt).items():
        if names is not None and name not in names:
            continue
        ln, _, _ = block_info(hdr, name, fmt)
        if ln == 0:
            continue
        a = np.zeros(ln * k, dtype=dt)
        rc = _abi.lib().psr_snapshot_read_block(fmt.encode(), str(path).encode(), C.byref(hdr), name.encode(),
                                                a.ctypes.data, err, 512)
        if rc:
            raise PsrError(rc, err.value.decode())
        out[name] = a.reshape(ln, 3) if k == 3 else a
    return hdr, out


def ye_of(ps, units):
    """psys_ye: free electrons per hydrogen nucleus from H and He
    (lambda_to_ye, src/psystem.h:134-138; yeMET is not part of it)"""
    xHeIII = 1.0 - ps.lambdaHeII - ps.lambdaHeI
    return (1.0 - ps.lambdaH) + units.C_HEMF / units.C_HMF * 0.25 * (ps.lambdaHeII + 2 * xHeIII)


def gas_blocks(path, fmt="psphray"):
    """the gas blocks an epoch reads (src/psystem.c:333-373: pos, mass, sml,
    rho, ie, ye, xHI, id), type 0's rows of each"""
    hdr, b = read(path, names=("pos", "id", "mass", "sml", "rho", "ie", "ye", "xHI"), fmt=fmt)
    n = int(hdr.N[0])
    return hdr, {k: v[:n] for k, v in b.items()}


def write_output(path, ps, ids, units, boxsize, redshift=0.0, time=1.0, src_pos=None, src_ngd=None):
    """a snapshot as psystem_write_output writes it (one file)"""
    n = int(ps.npar)
    ns = 0 if src_pos is None else len(src_ngd)
    hdr = make_header(n, ns, boxsize=boxsize, redshift=redshift, time=time, h=units.C_H)
    pos = np.zeros((n + ns, 3), np.float32)
    pos[:n] = ps.pos[:n]
    if ns:
        pos[n:] = np.asarray(src_pos, np.float64).reshape(-1, 3)
    idb = np.zeros(n + ns, np.uint64)
    idb[:n] = ids[:n]
    mass = np.zeros(n + ns, np.float32)
    mass[:n] = ps.mass[:n]
    blocks = {"pos": pos, "id": idb, "mass": mass, "sml": ps.sml[:n], "rho": ps.rho[:n],
              "xHI": ps.lambdaH[:n], "xHeI": ps.lambdaHeI[:n], "xHeII": ps.lambdaHeII[:n],
              "ye": ye_of(ps, units)[:n], "gammaHI": ps.yGrecHII[:n], "ie": ps.ie[:n],
              "lasthit": ps.lasthit[:n], "hits": ps.hits[:n]}
    if ns:
        blocks["ngammas"] = np.asarray(src_ngd, np.float64)
    write(path, hdr, blocks)


def write_output_device(path, ctx, ps, ids, units, boxsize, redshift=0.0, time=1.0, src_pos=None, src_ngd=None):
    """write_output with the blocks the ticks change gathered on the device
    (psr_output_gather); positions, sml, rho and mass, which an epoch keeps,
    from the host arrays ps"""
    n = int(ps.npar)
    g = ctx.output_gather(n)
    view = ParticleSystem(0)
    view.npar = n
    for f in ("pos", "sml", "rho", "mass"):
        setattr(view, f, getattr(ps, f)[:n])
    ns = 0 if src_pos is None else len(src_ngd)
    hdr = make_header(n, ns, boxsize=boxsize, redshift=redshift, time=time, h=units.C_H)
    pos = np.zeros((n + ns, 3), np.float32)
    pos[:n] = view.pos
    if ns:
        pos[n:] = np.asarray(src_pos, np.float64).reshape(-1, 3)
    idb = np.zeros(n + ns, np.uint64)
    idb[:n] = ids[:n]
    mass = np.zeros(n + ns, np.float32)
    mass[:n] = view.mass
    blocks = {"pos": pos, "id": idb, "mass": mass, "sml": view.sml, "rho": view.rho, "xHI": g["xHI"],
              "xHeI": g["xHeI"], "xHeII": g["xHeII"], "ye": g["ye"], "gammaHI": g["gammaHI"], "ie": g["ie"],
              "lasthit": g["lasthit"], "hits": g["hits"]}
    if ns:
        blocks["ngammas"] = np.asarray(src_ngd, np.float64)
    write(path, hdr, blocks)


_SNAP_FIELDS = (("id", np.uint64), ("pos", np.float32), ("mass", np.float32), ("sml", np.float32),
                ("rho", np.float32), ("ie", np.float32), ("xHI", np.float32), ("ye", np.float32))


def _epoch_snap(snap):
    """(psr_epoch_snap, the arrays it points into) for a dict of gas blocks"""
    keep = {k: np.ascontiguousarray(snap[k], dt) for k, dt in _SNAP_FIELDS if k in snap}
    es = _abi.EpochSnap(len(keep["id"]), *[keep[k].ctypes.data if k in keep else None for k, _ in _SNAP_FIELDS])
    return es, keep


def epoch_load(snap, units, h_only, npar_max=None):
    """(ParticleSystem, ids) of the first epoch from a snapshot's gas
    blocks (id, pos, mass, sml, rho, ie, xHI, ye): psr_epoch_load, with the
    reference's capacity of 1.05 ngas by default (src/psystem.c:302)"""
    n = len(snap["id"])
    cap = int(n * 1.05) if npar_max is None else int(npar_max)
    ps = ParticleSystem(cap)
    ids = np.zeros(cap, np.uint64)
    st = ps.as_struct()
    es, keep = _epoch_snap(snap)
    err = _err()
    rc = _abi.lib().psr_epoch_load(C.byref(st), ids.ctypes.data, cap, C.byref(es), C.byref(units), int(h_only),
                                   err, 512)
    if rc:
        raise PsrError(rc, err.value.decode())
    ps.npar = st.npar
    return ps, ids


def epoch_switch(ps, ids, snap, units, h_only, prev_duration, prev_nticks, tick_time, periodic=False,
                 boxsize=0.0):
    """psr_epoch_switch on ps (its arrays' length is the capacity npar_max,
    ps.npar in use) and ids; snap: dict with id, pos, mass, sml, rho, ie,
    xHI.  Returns the stats (matched, added, lost, removed, mean_shift);
    ps.npar is the new count."""
    cap = len(ps.mass)
    assert ids.dtype == np.uint64 and ids.flags.c_contiguous and len(ids) >= cap
    st = ps.as_struct()
    es, keep = _epoch_snap(snap)
    out = _abi.EpochStats()
    err = _err()
    rc = _abi.lib().psr_epoch_switch(C.byref(st), ids.ctypes.data, cap, C.byref(es), C.byref(units), int(h_only),
                                     int(periodic), float(boxsize), float(prev_duration), int(prev_nticks),
                                     float(tick_time), C.byref(out), err, 512)
    ps.npar = st.npar
    if rc:
        raise PsrError(rc, err.value.decode())
    return {k: getattr(out, k) for k in ("matched", "added", "lost", "removed", "mean_shift")}


def resized(ps, n):
    """a ParticleSystem of capacity n holding ps's first ps.npar particles"""
    out = ParticleSystem(n)
    for name in out.fields():
        getattr(out, name)[: ps.npar] = getattr(ps, name)[: ps.npar]
    out.npar = ps.npar
    return out
