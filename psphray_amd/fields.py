"""Synthetic SPH fields and sources for the five benchmark workloads.

The reference ships no generator (its Iliev test inputs are not in the
repository), so these definitions are this project's: jittered or
Zel'dovich-displaced lattices, lognormal densities, smoothing lengths that
enclose 32 neighbours, point sources.  Every quantity is in the reference's
internal units (src/units.c).  Seeds are 1000 + workload number.
"""
import numpy as np

from . import ParticleSystem, Physics, default_units, make_flags
from .rays import RayEmitter

KPC_BOX = 6.6      # box side, kpc (Iliev test 1)
NH_MEAN = 1e-3     # cm^-3
T_INIT = 1e4       # K
XHII_INIT = 1.2e-3
TICK_MYR = 0.05


def smoothing_lengths(pos, nngb=32, boxsize=None):
    """distance to the nngb-th nearest other particle (vacuum boundary)"""
    from scipy.spatial import cKDTree
    tree = cKDTree(pos)
    d, _ = tree.query(pos, k=nngb + 1, workers=-1)
    return d[:, nngb].astype(np.float32), d


def grf(n, slope, rng):
    """Gaussian random field on an n^3 mesh with P(k) ~ k^slope, unit variance"""
    k = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    amp[kk > 0] = kk[kk > 0] ** (slope / 2.0)
    noise = np.fft.rfftn(rng.standard_normal((n, n, n)))
    g = np.fft.irfftn(noise * amp, s=(n, n, n), axes=(0, 1, 2))
    g -= g.mean()
    return g / g.std()


def zeldovich_displacement(n, slope, rng):
    """displacement field psi = grad(phi) with delta(k) ~ k^slope/2 noise"""
    k = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    kx, ky, kzz = np.meshgrid(k, k, kz, indexing="ij")
    kk2 = kx ** 2 + ky ** 2 + kzz ** 2
    amp = np.zeros_like(kk2)
    amp[kk2 > 0] = np.sqrt(kk2[kk2 > 0]) ** (slope / 2.0)
    dk = np.fft.rfftn(rng.standard_normal((n, n, n))) * amp
    inv = np.zeros_like(kk2)
    inv[kk2 > 0] = 1.0 / kk2[kk2 > 0]
    psi = [np.fft.irfftn(1j * kv * inv * dk, s=(n, n, n), axes=(0, 1, 2)) for kv in (kx, ky, kzz)]
    return np.stack(psi, axis=-1)


class Workload:
    """particles + sources + spectrum + run parameters of one workload"""

    def __init__(self, name, ps, boxsize, physics, emitter, nray, tick_time, field_offsets=None,
                 boxsizes=None, per_field=None):
        self.name = name
        self.ps = ps
        self.boxsize = boxsize
        self.physics = physics
        self.emitter = emitter
        self.nray = nray
        self.tick_time = tick_time
        self.field_offsets = field_offsets
        self.boxsizes = boxsizes
        self.per_field = per_field

    def rays(self, tick):
        return self.emitter.emit(tick, self.nray, per_field=self.per_field)


def _init_state(ps, phys, T, xhii, he=False):
    u = phys.units
    ps.lambdaH[:] = 1.0 - xhii
    if he:
        ps.lambdaHeI[:] = 1.0 - xhii
        ps.lambdaHeII[:] = xhii
    else:
        ps.lambdaHeI[:] = 1.0
        ps.lambdaHeII[:] = 0.0
    ye = (1.0 - ps.lambdaH) + u.C_HEMF / u.C_HMF * 0.25 * (ps.lambdaHeII + 2 * (1.0 - ps.lambdaHeII - ps.lambdaHeI))
    ps.ie[:] = phys.Tye2ie(T, ye).astype(np.float32)


def _lattice(n, boxsize, rng, jitter=0.25):
    sp = boxsize / n
    g = (np.indices((n, n, n)).reshape(3, -1).T + 0.5)
    g = g + rng.uniform(-jitter, jitter, g.shape)
    return (g * sp), sp


def make_workload(cfg, rates_path, n=None, nray=None, seed=None, nfield=None, nsrc=None):
    """cfg in 1..5; n overrides the lattice size (tests use small fields)"""
    seed = 1000 + cfg if seed is None else seed
    rng = np.random.default_rng(seed)
    he = cfg == 3
    hmf = 0.76 if he else 1.0
    units = default_units(hmf=hmf)
    flags = make_flags(isothermal=not he, on_the_spot=True, h_only=not he,
                       secondary_ionization=True, comoving=False)
    phys = Physics(rates_path, units, flags, with_secion=he)
    u = units
    boxsize = KPC_BOX * u.U_KPC
    tick_time = TICK_MYR * u.U_MYR
    nH = NH_MEAN / u.U_CM ** 3                  # per internal volume
    rho_mean = nH * u.U_MPROTON / hmf
    per_sec = 1.0 / u.U_SEC                     # photons/s -> per internal time
    e136 = 13.6 * u.U_EV

    if cfg == 5:
        return _make_batched(rng, rates_path, phys, boxsize, tick_time, rho_mean, per_sec, e136,
                             nfield=nfield or 96, per_field=nray or 256)
    if cfg == 1:
        n = n or 32
        pos, sp = _lattice(n, boxsize, rng)
        N = len(pos)
        ps = ParticleSystem(N)
        ps.pos[:] = pos
        ps.sml[:], _ = smoothing_lengths(ps.pos.astype(np.float64))
        ps.rho[:] = rho_mean
        ps.mass[:] = rho_mean * boxsize ** 3 / N
        src_pos = np.zeros((1, 3))
        ngd = np.array([5e48 * per_sec])
        spectrum = ([e136], [1.0])
        nr = nray or 32
    elif cfg in (2, 4):
        n = n or 128
        pos, sp = _lattice(n, boxsize, rng)
        N = len(pos)
        ps = ParticleSystem(N)
        ps.pos[:] = pos
        ps.sml[:], _ = smoothing_lengths(ps.pos.astype(np.float64))
        g = grf(n, -2.5, rng).reshape(-1)
        rho = np.exp(1.0 * g)
        rho *= rho_mean / rho.mean()
        ps.rho[:] = rho
        ps.mass[:] = rho * boxsize ** 3 / N
        ns = nsrc or 64
        src_pos = rng.uniform(0, boxsize, (ns, 3))
        ngd = 10 ** rng.uniform(48, 50, ns) * per_sec
        spectrum = ([e136], [1.0])
        nr = nray or 4096
    elif cfg == 3:
        n = n or 256
        g0 = (np.indices((n, n, n)).reshape(3, -1).T + 0.5)
        psi = zeldovich_displacement(n, -2.0, rng).reshape(-1, 3)
        psi *= 0.3 / np.sqrt((psi ** 2).sum(axis=1).mean())
        sp = boxsize / n
        pos = np.mod((g0 + psi) * sp, boxsize)
        pos = np.minimum(pos, np.nextafter(boxsize, 0))
        N = len(pos)
        ps = ParticleSystem(N)
        ps.pos[:] = pos
        sml, d = smoothing_lengths(ps.pos.astype(np.float64))
        ps.sml[:] = sml
        m = rho_mean * boxsize ** 3 / N
        ps.mass[:] = m
        # SPH density from the 32 neighbours (cubic spline, support h)
        r = d / sml[:, None].astype(np.float64)
        w = np.where(r <= 0.5, 1 - 6 * r ** 2 + 6 * r ** 3, np.where(r <= 1, 2 * (1 - r) ** 3, 0.0))
        ps.rho[:] = (m * (8 / np.pi) * w.sum(axis=1) / sml.astype(np.float64) ** 3).astype(np.float32)
        ns = nsrc or 1024
        src_pos = rng.uniform(0, boxsize, (ns, 3))
        ngd = np.exp(rng.normal(np.log(1e49), 1.5, ns)) * per_sec
        spectrum = blackbody_bins(1e5, 16, u)
        nr = nray or 65536
    else:
        raise ValueError(f"unknown workload {cfg}")
    _init_state(ps, phys, T_INIT, XHII_INIT, he=he)
    em = RayEmitter(seed, src_pos, ngd, boxsize, tick_time, spectrum)
    return Workload(f"config{cfg}", ps, boxsize, phys, em, nr, tick_time)


def blackbody_bins(T, nbins, u):
    """photon-number weights of a blackbody in nbins log bins over 1-10 Ry
    (energies in internal units)"""
    edges = np.geomspace(1.0, 10.0, nbins + 1) * 13.6
    mid = np.sqrt(edges[1:] * edges[:-1])
    kT = 8.617333e-5 * T
    # photon number ~ E^2 / (exp(E/kT) - 1) dE
    w = mid ** 2 / np.expm1(mid / kT) * np.diff(edges)
    return (mid * u.U_EV, w / w.sum())


def _make_batched(rng, rates_path, phys, boxsize, tick_time, rho_mean, per_sec, e136, nfield,
                  per_field):
    sizes = rng.choice([8, 16, 24, 32, 48, 64], nfield)
    ps_list, src_pos, ngd, sfield = [], [], [], []
    for f, n in enumerate(sizes):
        pos, sp = _lattice(int(n), boxsize, rng)
        N = len(pos)
        ps = ParticleSystem(N)
        ps.pos[:] = pos
        ps.sml[:], _ = smoothing_lengths(ps.pos.astype(np.float64))
        g = grf(int(n), -2.5, rng).reshape(-1)
        rho = np.exp(0.8 * g)
        rho *= rho_mean / rho.mean()
        ps.rho[:] = rho
        ps.mass[:] = rho * boxsize ** 3 / N
        _init_state(ps, phys, T_INIT, XHII_INIT)
        ps_list.append(ps)
        k = int(rng.integers(1, 17))
        src_pos.append(rng.uniform(0, boxsize, (k, 3)))
        ngd.append(10 ** rng.uniform(47, 49, k) * per_sec)
        sfield.append(np.full(k, f, np.int32))
    N = sum(p.npar for p in ps_list)
    ps = ParticleSystem(N)
    off = np.zeros(nfield + 1, np.int64)
    off[1:] = np.cumsum([p.npar for p in ps_list])
    for f, p in enumerate(ps_list):
        for name in ps.fields():
            getattr(ps, name)[off[f]:off[f + 1]] = getattr(p, name)
    boxsizes = np.full(nfield, boxsize)
    em = RayEmitter(1005, np.concatenate(src_pos), np.concatenate(ngd), boxsize, tick_time,
                    ([e136], [1.0]), field=np.concatenate(sfield), boxsizes=boxsizes)
    return Workload("config5", ps, boxsize, phys, em, per_field * nfield, tick_time,
                    field_offsets=off, boxsizes=boxsizes, per_field=per_field)
