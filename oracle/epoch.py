"""Oracle (test infrastructure only): a CPU restatement of the reference's
snapshot format and epoch switch, used by the tests to check the library's
host code (psr_snapshot_*, psr_epoch_load, psr_epoch_switch).  Nothing in
the product imports it.

- the psphray snapshot format (src/reader-psphray.inc): the header record
  (struct header_t, :1-28, its first 256 bytes), block order (_blockname,
  :55-66), item sizes (_itemsize, :68-90), lengths (_length, :115-330),
  offsets (_offset, :332-344), record framing (_read / _write, :400-460);
  written and parsed here with numpy, from this table alone;
- psystem_read_epoch (src/psystem.c:300-379), psystem_match_epoch
  (:381-504) with the ID hash chains of idhash_rebuild (:49-58), dist
  (:507-518), compress (:109-139) and the lasthit rescale (:266-272), as
  Python loops over the same chains the reference walks (small inputs).

Pins: the header offsets are checked against the C struct's natural layout
(boxsize at 128, OmegaB at 200, nrays at 208, flags at 216), the block
table against hand-counted file sizes, and the epoch switch against
invariants (identity on an unchanged snapshot, every ID kept once, matched
state carried, unmatched state reset) in the tests.
"""
import math

import numpy as np

HEADER_DTYPE = np.dtype({
    "names": ["N", "mass", "time", "redshift", "flag_sfr", "flag_feedback", "Nparticle_total_low",
              "flag_cool", "Nfiles", "boxsize", "OmegaM", "OmegaL", "h", "flag_sft", "flag_met",
              "Nparticle_total_high", "flag_entropy", "OmegaB", "nrays", "flag_Hmf", "flag_Hemf",
              "flag_Helium", "flag_gammaHI", "flag_cloudy", "flag_eos"],
    "formats": [("<u4", 6), ("<f8", 6), "<f8", "<f8", "<i4", "<i4", ("<u4", 6), "<i4", "<i4", "<f8", "<f8", "<f8",
                "<f8", "<i4", "<i4", ("<u4", 6), "<i4", "<f8", "<u8", "<i4", "<i4", "<i4", "<i4", "<i4", "<i4"],
    "offsets": [0, 24, 72, 80, 88, 92, 96, 120, 124, 128, 136, 144, 152, 160, 164, 168, 192, 200, 208, 216,
                220, 224, 228, 232, 236],
    "itemsize": 256,
})

# (name, item bytes, numpy type, components, particle types counted, flag)
BLOCKS = [
    ("pos", 12, "<f4", 3, (0, 4, 5), None), ("id", 8, "<u8", 1, (0, 4, 5), None),
    ("mass", 4, "<f4", 1, (0, 4, 5), None), ("ie", 4, "<f4", 1, (0,), None), ("rho", 4, "<f4", 1, (0,), None),
    ("ye", 4, "<f4", 1, (0,), None), ("xHI", 4, "<f4", 1, (0,), None), ("sml", 4, "<f4", 1, (0,), None),
    ("Hmf", 4, "<f4", 1, (0,), "flag_Hmf"), ("Hemf", 4, "<f4", 1, (0,), "flag_Hemf"),
    ("xHeI", 4, "<f4", 1, (0,), "flag_Helium"), ("xHeII", 4, "<f4", 1, (0,), "flag_Helium"),
    ("gammaHI", 4, "<f4", 1, (0,), "flag_gammaHI"), ("HIa_cloudy", 4, "<f4", 1, (0,), "flag_cloudy"),
    ("eos", 4, "<f4", 1, (0,), "flag_eos"), ("lasthit", 8, "<i8", 1, (0,), None),
    ("hits", 4, "<i4", 1, (0,), None), ("sft", 4, "<f4", 1, (4,), None), ("ngammas", 8, "<f8", 1, (5,), None),
    ("spec", 8, "S8", 1, (5,), None),
]
# src/reader-massiveblack.inc:26-330
BLOCKS_MASSIVEBLACK = [
    ("pos", 12, "<f4", 3, (0, 1, 4, 5), None), ("vel", 12, "<f4", 3, (0, 1, 4, 5), None),
    ("id", 8, "<u8", 1, (0, 1, 4, 5), None), ("mass", 4, "<f4", 1, (0, 4, 5), None),
    ("ie", 4, "<f4", 1, (0,), None), ("rho", 4, "<f4", 1, (0,), None), ("ye", 4, "<f4", 1, (0,), None),
    ("xHI", 4, "<f4", 1, (0,), None), ("sml", 4, "<f4", 1, (0,), None), ("sfr", 4, "<f4", 1, (0,), "flag_sfr"),
    ("sft", 4, "<f4", 1, (4,), "flag_sft"), ("met", 4, "<f4", 1, (0, 4), "flag_met"),
    ("bhmass", 4, "<f4", 1, (5,), None), ("bhmdot", 4, "<f4", 1, (5,), None),
]
FORMATS = {"psphray": BLOCKS, "massiveblack": BLOCKS_MASSIVEBLACK}


def block_length(hdr, types, flag):
    if flag is not None and not hdr[flag]:
        return 0
    return sum(int(hdr["N"][t]) for t in types)


def write_snapshot(path, hdr, blocks, fmt="psphray"):
    """header record, then each block the header calls for, in order,
    framed by its byte count; absent blocks are zeros"""
    with open(path, "wb") as f:
        np.array([256], "<i4").tofile(f)
        rec = np.zeros(1, HEADER_DTYPE)          # zeros in the alignment gaps
        rec[0] = hdr
        rec.tofile(f)
        np.array([256], "<i4").tofile(f)
        for name, item, dt, k, types, flag in FORMATS[fmt]:
            ln = block_length(hdr, types, flag)
            if ln == 0:
                continue
            a = np.zeros(ln * k, dt) if name not in blocks else np.asarray(blocks[name], dt).reshape(-1)
            assert a.size == ln * k and a.nbytes == ln * item, name
            np.array([ln * item], "<i4").tofile(f)
            a.tofile(f)
            np.array([ln * item], "<i4").tofile(f)


def read_snapshot(path, fmt="psphray"):
    raw = np.fromfile(path, np.uint8)
    assert raw[:4].view("<i4")[0] == 256 and raw[260:264].view("<i4")[0] == 256
    hdr = raw[4:260].view(HEADER_DTYPE)[0]
    out, off = {}, 264
    for name, item, dt, k, types, flag in FORMATS[fmt]:
        ln = block_length(hdr, types, flag)
        if ln == 0:
            continue
        nb = ln * item
        assert raw[off:off + 4].view("<i4")[0] == nb
        a = raw[off + 4:off + 4 + nb].view(dt).copy()
        assert raw[off + 4 + nb:off + 8 + nb].view("<i4")[0] == nb
        out[name] = a.reshape(ln, 3) if k == 3 else a
        off += nb + 8
    assert off == len(raw)
    return hdr, out


# ------------------------------------------------------------------ epochs
IDHASHMASK = (1 << 28) - 1
PF_INVALID = 1


def _f32(x):
    return float(np.float32(x))


def _fdim(a, b):
    return a - b if a > b else 0.0


def _set_fraction(x, rest):
    """psys_set_lambdaH / psys_set_lambdaHe (src/psystem.h:156-169)"""
    if x >= 1.0 or rest <= 0.0:
        return 1.0
    if x <= 0.0 or rest >= 1.0:
        return 0.0
    return x


class Psys:
    """the reference's PSystem arrays (src/psystem.h:32-78), capacity npar_max"""
    F32 = ("sml", "mass", "rho", "ie", "yeMET", "yGdepHI", "yGdepHeI", "yGdepHeII", "yGrecHII", "yGrecHeII",
           "yGrecHeIII", "heat")
    F64 = ("lambdaH", "lambdaHeI", "lambdaHeII")

    def __init__(self, npar_max):
        self.npar = 0
        self.npar_max = npar_max
        self.pos = np.zeros((npar_max, 3), np.float32)
        for f in self.F32:
            setattr(self, f, np.zeros(npar_max, np.float32))
        for f in self.F64:
            setattr(self, f, np.zeros(npar_max, np.float64))
        self.flag = np.zeros(npar_max, np.int8)
        self.lasthit = np.zeros(npar_max, np.int64)
        self.hits = np.zeros(npar_max, np.int32)
        self.id = np.zeros(npar_max, np.uint64)

    def swap(self, i, j):
        for f in ("pos", "flag", "lasthit", "hits", "id") + self.F32 + self.F64:
            a = getattr(self, f)
            a[[i, j]] = a[[j, i]]


def read_epoch(ps, snap, C_HMF, h_only):
    """psystem_read_epoch (src/psystem.c:300-379) for one file"""
    n = len(snap["id"])
    for i in range(n):
        ps.pos[i] = snap["pos"][i]
        for f in ("mass", "sml", "rho", "ie"):
            getattr(ps, f)[i] = snap[f][i]
        ps.id[i] = snap["id"][i]
        xHI = float(snap["xHI"][i])
        xHII = 1.0 - xHI
        ps.lambdaH[i] = _set_fraction(xHI, xHII)
        yeMET = _fdim(float(snap["ye"][i]), xHII)
        if not h_only:
            xHeII = yeMET * C_HMF / (1 - C_HMF) * 4
            xHeIII = _fdim(xHeII, 1.0)
            xHeII = min(1.0, xHeII)
            xHeIII = min(1.0 - xHeII, xHeIII)
            xHeI = 1.0 - xHeII - xHeIII
            ps.lambdaHeI[i] = _set_fraction(xHeI, xHeII + xHeIII)
            ps.lambdaHeII[i] = _set_fraction(xHeII, xHeI + xHeIII)
            yeMET = _fdim(yeMET, (xHeII + 2 * xHeIII) * (1 - C_HMF) / C_HMF * 0.25)
        ps.yeMET[i] = yeMET
    ps.npar = n


def dist(p1, p2, periodic, boxsize):
    """src/psystem.c:507-518: float differences, float squares, double sum"""
    r = 0.0
    for d in range(3):
        dd = abs(np.float32(p1[d]) - np.float32(p2[d]))
        if periodic and float(dd) > 0.5 * boxsize:
            dd = np.float32(boxsize - float(dd))
        r += float(dd * dd)
    return _f32(math.sqrt(r))


def match_epoch(ps, snap, C_HMF, h_only, periodic=False, boxsize=0.0):
    """psystem_match_epoch (src/psystem.c:381-504) for one file; returns
    (new, lost, distsum).  A particle unmatched when the store is full is
    dropped (the reference writes it one past npar_max)."""
    head = {}
    nxt = np.full(ps.npar, -1, np.int64)
    for i in range(ps.npar):            # idhash_rebuild
        h = int(ps.id[i]) & IDHASHMASK
        nxt[i] = head.get(h, -1)
        head[h] = i
    ps.flag[:ps.npar] |= PF_INVALID
    new = lost = 0
    distsum = 0.0
    for ip in range(len(snap["id"])):
        h = int(snap["id"][ip]) & IDHASHMASK
        best, mindist = -1, _f32(500.0)
        j = head.get(h, -1)
        while j != -1:
            dst = dist(snap["pos"][ip], ps.pos[j], periodic, boxsize)
            if dst < mindist:
                mindist, best = dst, j
            j = nxt[j]
        insert = best == -1
        if insert:
            if ps.npar == ps.npar_max:
                lost += 1
                continue
            best = ps.npar
            ps.npar += 1
            new += 1
        else:
            distsum += mindist
        ps.pos[best] = snap["pos"][ip]
        for f in ("mass", "sml", "rho", "ie"):
            getattr(ps, f)[best] = snap[f][ip]
        ps.flag[best] &= ~PF_INVALID
        if insert:
            ps.id[best] = snap["id"][ip]
            xHI = float(snap["xHI"][ip])
            xHII = 1.0 - xHI
            ps.lambdaH[best] = _set_fraction(xHI, xHII)
            yeMET = _fdim(float(ps.yeMET[best]), xHII)
            if not h_only:
                xHeII = yeMET * C_HMF / (1 - C_HMF) * 4
                xHeIII = min(1.0, xHeII)       # the reference overwrites fdim(xHeII, 1) so
                xHeII = min(1.0, xHeII)
                xHeI = 1.0 - xHeII - xHeIII
                ps.lambdaHeI[best] = _set_fraction(xHeI, xHeII + xHeIII)
                ps.lambdaHeII[best] = _set_fraction(xHeII, xHeI + xHeIII)
                yeMET = _fdim(yeMET, (xHeII + xHeIII) * (1 - C_HMF) / C_HMF * 0.25)
            ps.yeMET[best] = yeMET
            for f in ("yGdepHI", "yGdepHeI", "yGdepHeII", "yGrecHII", "yGrecHeII", "yGrecHeIII"):
                getattr(ps, f)[best] = 0.0
            if ps.npar_max > 3:                # the loop counter i is 3 here (src/psystem.c:483-484)
                ps.heat[3] = 0.0
                ps.hits[3] = 0
            ps.lasthit[best] = 0
            ps.flag[best] = 0
    return new, lost, distsum


def rescale_lasthit(ps, prev_duration, prev_nticks, tick_time):
    """src/psystem.c:266-272: lasthit in ticks of the new epoch (C truncation)"""
    last_ticktime = prev_duration / prev_nticks
    for i in range(ps.npar):
        ps.lasthit[i] = int((int(ps.lasthit[i]) * last_ticktime - prev_duration) / tick_time)


def compress(ps):
    """src/psystem.c:109-139: the last valid particles fill the holes"""
    j = ps.npar - 1
    while j > 0 and ps.flag[j] & PF_INVALID:
        j -= 1
    i = 0
    while i < j:
        if ps.flag[i] & PF_INVALID:
            ps.swap(i, j)
            while ps.flag[j] & PF_INVALID and j > i:
                j -= 1
        i += 1
    if ps.npar > 0 and j == 0 and ps.flag[0] & PF_INVALID:
        j = -1                                # every particle removed (the reference aborts)
    ps.npar = j + 1


def switch_epoch(ps, snap, C_HMF, h_only, prev_duration, prev_nticks, tick_time, periodic=False, boxsize=0.0):
    """match, lasthit rescale, compress (src/psystem.c:259-275)"""
    new, lost, distsum = match_epoch(ps, snap, C_HMF, h_only, periodic, boxsize)
    rescale_lasthit(ps, prev_duration, prev_nticks, tick_time)
    removed = int(((ps.flag[:ps.npar] & PF_INVALID) != 0).sum())
    compress(ps)
    return dict(added=new, lost=lost, removed=removed, distsum=distsum)
