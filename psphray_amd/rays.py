"""Counter-based ray lists (Philox4x32-10, Salmon et al. 2011).

Replaces the GSL streams of emit_rays (src/loop.c:163-373) with counter-based
streams keyed by (seed, tick, ray), so the GPU path and the CPU oracle consume
identical rays.  The list of a tick is a function of those keys and of each
source's lastemit tick (src/sources.c:67-78), which emit() advances like the
reference: a RayEmitter must emit ticks 1, 2, 3, ... in order, once each
(a fresh emitter replays the same sequence).

The steps follow the reference (src/loop.c:163-373):
- the source of each ray is drawn with weight Ngamma_dot * (tick - lastemit)
  * tick_time (src/sources.c:67-78) from word 0 of the ray's block 0;
- rays are ordered by source, stably (the reference's qsort by (type, isrc),
  src/loop.c:236), and a source's rays share its weight equally
  (src/loop.c:305-321);
- the direction is isotropic, drawn as gsl_ran_dir_3d draws it (src/loop.c:250:
  x, y uniform in [-1, 1] until s = x^2 + y^2 <= 1, then z = -1 + 2 s and
  (x, y) scaled by 2 sqrt(1 - s); Knop, CACM 13, 326, 1970) from blocks
  1, 2, ... of the ray's stream, two tries per block;
- the photon energy is a bin of the source's spectrum drawn from word 3 of
  block 0 (src/spectra.c:50-60);
- the length is 2 * boxsize (src/loop.c:267).
Every operation is a correctly rounded +, *, / or sqrt, so the device and the
oracle reproduce the list bit for bit.
"""
import numpy as np

from ._abi import RAY_DTYPE

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
KEY1 = 0x50535248        # second key word of the emission streams
DIR_BLOCKS = 32          # blocks of direction tries (64 tries) before the fallback (0, 0, 1)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """vectorised Philox4x32-R on uint32 arrays; returns four uint32 arrays"""
    c0, c1, c2, c3 = (np.asarray(x, np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(rounds):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + _W0)
            k1 = np.uint32(k1 + _W1)
    return c0, c1, c2, c3


def uniform01(x):
    """uint32 -> double in (0, 1)"""
    return (x.astype(np.float64) + 0.5) * (1.0 / 4294967296.0)


def isotropic_dirs(ids, tick, seed):
    """unit vectors (float64 x, y, z) for ray ids: the rejection method of
    gsl_ran_dir_3d on blocks 1..DIR_BLOCKS of each ray's stream"""
    n = len(ids)
    x = np.zeros(n)
    y = np.zeros(n)
    z = np.ones(n)
    done = np.zeros(n, bool)
    lo = np.full(n, tick & 0xFFFFFFFF, np.uint32)
    hi = np.full(n, tick >> 32, np.uint32)
    for blk in range(1, DIR_BLOCKS + 1):
        todo = np.nonzero(~done)[0]
        if len(todo) == 0:
            break
        w = philox4x32(ids[todo], lo[todo], hi[todo], np.full(len(todo), blk, np.uint32), seed, KEY1)
        for a, b in ((w[0], w[1]), (w[2], w[3])):
            xx = -1.0 + 2.0 * uniform01(a)
            yy = -1.0 + 2.0 * uniform01(b)
            s = xx * xx + yy * yy
            ok = (s <= 1.0) & ~done[todo]
            sel = todo[ok]
            sc = 2.0 * np.sqrt(1.0 - s[ok])
            x[sel] = xx[ok] * sc
            y[sel] = yy[ok] * sc
            z[sel] = -1.0 + 2.0 * s[ok]
            done[sel] = True
    return x, y, z


def spectrum_cdf(weights):
    """cumulative weights normalised by their (sequential) sum"""
    c = np.cumsum(np.asarray(weights, np.float64))
    return c / c[-1]


class RayEmitter:
    """Emits the ray list of each tick for a set of point sources.

    sources: pos (ns,3) float64, ngamma_dot (ns,) photons per internal time,
    field (ns,) int (batched fields, default 0); spectrum: (energies,
    weights) in internal energy units (one bin = monochromatic), or a list of
    such spectra with spec (ns,) naming each source's."""

    def __init__(self, seed, pos, ngamma_dot, boxsize, tick_time, spectrum, field=None,
                 boxsizes=None, spec=None):
        self.seed = int(seed) & 0xFFFFFFFF
        self.pos = np.asarray(pos, np.float64).reshape(-1, 3)
        self.ngd = np.asarray(ngamma_dot, np.float64)
        self.ns = len(self.ngd)
        self.field = np.zeros(self.ns, np.int32) if field is None else np.asarray(field, np.int32)
        if boxsizes is None:
            boxsizes = np.full(int(self.field.max()) + 1, boxsize, np.float64)
        self.boxsizes = np.asarray(boxsizes, np.float64)
        self.tick_time = float(tick_time)
        spectra = spectrum if isinstance(spectrum, list) else [spectrum]
        self.spectra = [(np.atleast_1d(np.asarray(e, np.float64)), spectrum_cdf(np.atleast_1d(w)))
                        for e, w in spectra]
        self.spec = np.zeros(self.ns, np.int32) if spec is None else np.asarray(spec, np.int32)
        self.lastemit = np.zeros(self.ns, np.int64)

    def weights(self, tick):
        """Ngamma_dot * (tick - lastemit) * tick_time (src/sources.c:67-78)"""
        return self.ngd * (int(tick) - self.lastemit) * self.tick_time

    def groups(self, nray, per_field=None):
        """(sources, ray count, first ray id) per emission group: one group of
        every source, or one per field with per_field rays"""
        if per_field is None:
            return [(np.arange(self.ns), int(nray), 0)]
        nf = len(self.boxsizes)
        return [(np.nonzero(self.field == f)[0], int(per_field), f * int(per_field)) for f in range(nf)]

    def emit(self, tick, nray, per_field=None):
        """rays of one tick.  per_field: rays per field (batched runs), else
        nray rays over all sources."""
        tick = int(tick)
        out = []
        for srcs, nr, ray0 in self.groups(nray, per_field):
            if nr == 0 or len(srcs) == 0:
                continue
            ids = np.arange(nr, dtype=np.uint32) + np.uint32(ray0)
            r0, _, _, r3 = philox4x32(ids, np.full(nr, tick & 0xFFFFFFFF, np.uint32),
                                      np.full(nr, tick >> 32, np.uint32),
                                      np.zeros(nr, np.uint32), self.seed, KEY1)
            wts = self.weights(tick)[srcs]
            cdf = np.cumsum(wts)
            if cdf[-1] <= 0:
                continue
            pick = np.searchsorted(cdf / cdf[-1], uniform01(r0), side="right")
            pick = np.minimum(pick, len(srcs) - 1)
            isrc = srcs[pick]
            order = np.argsort(isrc, kind="stable")  # qsort by (type, isrc), src/loop.c:236
            isrc = isrc[order]
            ids = ids[order]
            r3 = r3[order]
            cnt = np.bincount(isrc, minlength=self.ns)
            rays = np.zeros(nr, RAY_DTYPE)
            dx, dy, dz = isotropic_dirs(ids, tick, self.seed)
            rays["dir"][:, 0] = dx.astype(np.float32)
            rays["dir"][:, 1] = dy.astype(np.float32)
            rays["dir"][:, 2] = dz.astype(np.float32)
            rays["s"] = self.pos[isrc].astype(np.float32)
            w_all = self.weights(tick)
            rays["Nph"] = w_all[isrc] * (1.0 / cnt[isrc])
            u3 = uniform01(r3)
            freq = np.empty(nr)
            for k in np.unique(self.spec[isrc]):
                eng, scdf = self.spectra[k]
                m = self.spec[isrc] == k
                ib = np.searchsorted(scdf, u3[m], side="right")
                freq[m] = eng[np.minimum(ib, len(eng) - 1)]
            rays["freq"] = freq
            fld = self.field[isrc]
            rays["length"] = self.boxsizes[fld] * 2
            rays["type"] = -1
            rays["field"] = fld
            rays["src_or_par"] = isrc
            self.lastemit[np.unique(isrc)] = tick
            out.append(rays)
        if not out:
            return np.zeros(0, RAY_DTYPE)
        return np.concatenate(out)
