"""Counter-based ray lists (Philox4x32-10, Salmon et al. 2011).

Replaces the GSL streams of emit_rays (src/loop.c:163-373) with counter-based
streams keyed by (seed, tick, ray), so the GPU path and the CPU oracle consume
identical rays.  The list of a tick is a function of those keys and of each
source's lastemit tick (src/sources.c:67-78), which emit() advances like the
reference: a RayEmitter must emit ticks 1, 2, 3, ... in order, once each
(a fresh emitter replays the same sequence).  The source choice follows the reference's weights
Ngamma_dot * (tick - lastemit) * tick_time (src/sources.c:67-78), rays from
one source share its weight equally (src/loop.c:305-321), directions are
isotropic and the length is 2 * boxsize (src/loop.c:267).
"""
import numpy as np

from ._abi import RAY_DTYPE

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


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


class RayEmitter:
    """Emits the ray list of each tick for a set of point sources.

    sources: pos (ns,3) float64, ngamma_dot (ns,) photons per internal time,
    field (ns,) int (batched fields, default 0); spectrum: (energies, weights)
    in internal energy units (one bin = monochromatic)."""

    def __init__(self, seed, pos, ngamma_dot, boxsize, tick_time, spectrum, field=None,
                 boxsizes=None):
        self.seed = int(seed) & 0xFFFFFFFF
        self.pos = np.asarray(pos, np.float64).reshape(-1, 3)
        self.ngd = np.asarray(ngamma_dot, np.float64)
        self.ns = len(self.ngd)
        self.field = np.zeros(self.ns, np.int32) if field is None else np.asarray(field, np.int32)
        if boxsizes is None:
            boxsizes = np.full(int(self.field.max()) + 1, boxsize, np.float64)
        self.boxsizes = np.asarray(boxsizes, np.float64)
        self.tick_time = float(tick_time)
        eng, w = spectrum
        self.eng = np.atleast_1d(np.asarray(eng, np.float64))
        w = np.atleast_1d(np.asarray(w, np.float64))
        self.spec_cdf = np.cumsum(w) / w.sum()
        self.lastemit = np.zeros(self.ns, np.int64)

    def emit(self, tick, nray, per_field=None):
        """rays of one tick.  per_field: rays per field (batched runs), else
        nray rays over all sources."""
        tick = int(tick)
        if per_field is None:
            groups = [(np.arange(self.ns), int(nray), 0)]
        else:
            nf = len(self.boxsizes)
            groups = [(np.nonzero(self.field == f)[0], int(per_field), f * int(per_field))
                      for f in range(nf)]
        out = []
        for srcs, nr, ray0 in groups:
            if nr == 0 or len(srcs) == 0:
                continue
            ids = np.arange(nr, dtype=np.uint32) + np.uint32(ray0)
            r0, r1, r2, r3 = philox4x32(ids, np.full(nr, tick & 0xFFFFFFFF, np.uint32),
                                        np.full(nr, tick >> 32, np.uint32),
                                        np.zeros(nr, np.uint32), self.seed, 0x50535248)
            wts = self.ngd[srcs] * (tick - self.lastemit[srcs]) * self.tick_time
            cdf = np.cumsum(wts)
            if cdf[-1] <= 0:
                continue
            pick = np.searchsorted(cdf / cdf[-1], uniform01(r0), side="right")
            pick = np.minimum(pick, len(srcs) - 1)
            isrc = srcs[pick]
            order = np.argsort(isrc, kind="stable")  # qsort by (type, isrc), src/loop.c:236
            isrc = isrc[order]
            r1, r2, r3 = r1[order], r2[order], r3[order]
            cnt = np.bincount(isrc, minlength=self.ns)
            rays = np.zeros(nr, RAY_DTYPE)
            ct = 2.0 * uniform01(r1) - 1.0
            phi = 2.0 * np.pi * uniform01(r2)
            st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
            rays["dir"][:, 0] = (st * np.cos(phi)).astype(np.float32)
            rays["dir"][:, 1] = (st * np.sin(phi)).astype(np.float32)
            rays["dir"][:, 2] = ct.astype(np.float32)
            rays["s"] = self.pos[isrc].astype(np.float32)
            w_all = self.ngd * (tick - self.lastemit) * self.tick_time
            rays["Nph"] = w_all[isrc] * (1.0 / cnt[isrc])
            ib = np.searchsorted(self.spec_cdf, uniform01(r3), side="right")
            rays["freq"] = self.eng[np.minimum(ib, len(self.eng) - 1)]
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
