"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden")

_D = C.c_double
_I32 = C.c_int32
_I64 = C.c_int64
_VP = C.c_void_p


class OrUnits(C.Structure):
    _fields_ = [(n, _D) for n in (
        "C_H C_HMF C_HEMF C_1_CMH C_1_GRAMH C_1_SECH "
        "U_CM U_GRAM U_SEC U_M U_KM U_KPC U_YR U_MYR U_KG U_MSUN U_MPROTON U_J U_EV U_ERG U_RY_ENG "
        "C_BOLTZMANN C_PLANCK C_SPEED_LIGHT C_H_PER_MASS C_HE_PER_MASS C_HI_ENERGY C_HEI_ENERGY "
        "C_HEII_ENERGY U_MPROTON_OVER_C_BOLTZMANN C_BOLTZMANN_OVER_U_MPROTON").split()]


class OrFlags(C.Structure):
    _fields_ = [(n, _I32) for n in
                "isothermal adiabatic on_the_spot h_only secondary_ionization trace_only "
                "no_photon comoving".split()] + [("redshift", _D), ("fake_temperature", _D)]


class OrPsys(C.Structure):
    _fields_ = [("npar", _I64)] + [(n, _VP) for n in (
        "pos sml flag mass rho ie yeMET lambdaH lambdaHeI lambdaHeII yGdepHI yGdepHeI yGdepHeII "
        "yGrecHII yGrecHeII yGrecHeIII heat lasthit hits").split()]


class OrStats(C.Structure):
    _fields_ = [(n, _I64) for n in
                "total_deposit_count saturated_deposit_count active_count evolve_count "
                "solver_steps".split()] + \
               [(n, _D) for n in
                "lost_photon_count_sum first_ionization_HI first_ionization_HeI "
                "first_ionization_HeII secondary_ionization_HI secondary_ionization_HeI "
                "trace_time deposit_time merge_time update_time".split()]


HIT_DTYPE = np.dtype([("ipar", "<i8"), ("d", "<f8"), ("b", "<f8")])
RAY_DTYPE = np.dtype([("s", "<f4", 3), ("dir", "<f4", 3), ("Nph", "<f8"), ("freq", "<f8"),
                      ("length", "<f8"), ("type", "<i4"), ("field", "<i4"),
                      ("src_or_par", "<i8")], align=True)
STEP_DTYPE = np.dtype([(n, "<f8") for n in
                       "lambdaH lambdaHeI lambdaHeII ie yeMET yGdepHI yGdepHeI yGdepHeII heat "
                       "time nH T yGrecHII yGrecHeII yGrecHeIII step_remain".split()] +
                      [("refined", "<i4"), ("nsteps", "<i4"), ("ipar", "<i8")], align=True)

_SIG = {
    "or_units_init": (None, [C.POINTER(OrUnits), _D, _D, _D, _D, _D, _D]),
    "or_create": (_VP, [C.POINTER(OrUnits), C.POINTER(OrFlags), C.c_char_p, _VP, _I32]),
    "or_destroy": (None, [_VP]),
    "or_error": (C.c_char_p, [_VP]),
    "or_ar_get": (_D, [_VP, _I32, _D]),
    "or_xs_get": (_D, [_VP, _I32, _D]),
    "or_sph_depth": (_D, [_VP, _D]),
    "or_secion_get": (_D, [_VP, _I32, _D, _D]),
    "or_ar_nrows": (_I32, [_VP]),
    "or_ar_table": (_I32, [_VP, _VP, C.POINTER(_D), C.POINTER(_D)]),
    "or_verner_hi": (C.c_float, [C.c_float]),
    "or_osterbrok_hei": (C.c_float, [C.c_float]),
    "or_osterbrok_heii": (C.c_float, [C.c_float]),
    "or_pluecker": (_I32, [_VP, _VP, _VP, _VP]),
    "or_tree_build": (_VP, [_VP, _VP, _I64, _D]),
    "or_tree_free": (None, [_VP]),
    "or_tree_ncells": (_I64, [_VP]),
    "or_tree_export": (None, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "or_tree_leaf_of": (None, [_VP, _VP]),
    "or_trace_rays": (_I64, [_VP, C.POINTER(OrPsys), _VP, _I64, _VP, _VP, _I64, _I32]),
    "or_deposit": (_I32, [_VP, C.POINTER(OrPsys), _VP, _I64, _VP, _VP, _VP, _VP,
                          C.POINTER(OrStats), _I32]),
    "or_merge": (_I64, [C.POINTER(OrPsys), _I64, _VP, _VP, _VP]),
    "or_update": (_I32, [_VP, C.POINTER(OrPsys), _VP, _I64, _I64, _D, C.POINTER(OrStats), _I32]),
    "or_tick": (_I32, [_VP, _VP, C.POINTER(OrPsys), _VP, _I64, _I64, _D, C.POINTER(OrStats), _I32]),
    "or_chem_rhs": (C.c_int, [_D, _VP, _VP, _VP]),
    "or_chem_jac": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "or_step_evolve": (_I32, [_VP, _VP, _D, _VP]),
    "or_riccati_tick": (_D, [_D, _D, _D, _D, _D]),
    "or_step_evolve_many": (_I32, [_VP, _VP, _I64, _D, _I32]),
    "or_ros3_fixed": (_I32, [_VP, _VP, _I32, _VP, _VP]),
    "or_ieye2T": (_D, [_VP, _D, _D]),
    "or_Tye2ie": (_D, [_VP, _D, _D]),
    "or_set_flags": (None, [_VP, C.POINTER(OrFlags)]),
    "or_emit_rays": (_I64, [_I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _I32, _VP, _I64, _D, _I64,
                            _I64, C.c_uint32, _VP, _I64]),
    "or_lte_tables": (None, [_VP, _VP, _VP, _VP]),
    "or_emit_recomb": (_I64, [_VP, _VP, C.POINTER(OrPsys), _VP, _I64, _I64, _I64, _D, _I32, _VP, _VP, _VP, _D,
                              _I64, C.c_uint32, _VP, _I64]),
}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref/libref.so built from the reference's own sources, or None"""
    if not os.path.exists(REF_LIB):
        return None
    L = C.CDLL(REF_LIB)
    L.pluecker_.restype = C.c_int
    L.pluecker_.argtypes = [_VP, _VP, _VP, _VP]
    L.sph_depth.restype = _D
    L.sph_depth.argtypes = [_D]
    L.ref_lsoda_evolve.restype = C.c_int
    L.ref_lsoda_evolve.argtypes = [_VP, _VP, _VP, C.c_int]
    L.ref_robertson_run.restype = C.c_int
    L.ref_robertson_run.argtypes = [_VP]
    return L


def p(a):
    return None if a is None else a.ctypes.data


def units(h=0.72, hmf=0.76, hemf=None):
    u = OrUnits()
    lib().or_units_init(C.byref(u), h, hmf, 1.0 - hmf if hemf is None else hemf,
                        3.0835455558318480e+21, 1.9847005219450705e+43, 3.08568025e+16)
    return u


def flags(isothermal=True, adiabatic=False, on_the_spot=True, h_only=True,
          secondary_ionization=True, trace_only=False, no_photon=False, comoving=False,
          redshift=0.0, fake_temperature=-1.0):
    return OrFlags(int(isothermal), int(adiabatic), int(on_the_spot), int(h_only),
                   int(secondary_ionization), int(trace_only), int(no_photon), int(comoving),
                   float(redshift), float(fake_temperature))


def reference_kernel_table():
    """src/kernel.akline values, kept as a fixture under tests/golden"""
    return np.load(os.path.join(GOLDEN, "kernel_akline.npy"))


def rates_path(name="Hui"):
    return os.path.join(GOLDEN, f"atomic_rates_{name}.txt")


class Oracle:
    def __init__(self, u=None, f=None, rates=None, kernel=None):
        self.L = lib()
        self.u = u if u is not None else units()
        self.f = f if f is not None else flags()
        k = np.ascontiguousarray(reference_kernel_table() if kernel is None else kernel, np.float64)
        self._k = k
        self.h = self.L.or_create(C.byref(self.u), C.byref(self.f),
                                  (rates or rates_path()).encode(), p(k), len(k))
        err = self.L.or_error(self.h)
        if err:
            raise RuntimeError(err.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_destroy(self.h)
            self.h = None

    def set_flags(self, f):
        self.f = f
        self.L.or_set_flags(self.h, C.byref(f))

    def lte_tables(self):
        """(energies [3][1024], cdf [3][128][1024], row bounds in T [127])"""
        eng = np.zeros((3, 1024))
        cdf = np.zeros((3, 128, 1024))
        tb = np.zeros(127)
        self.L.or_lte_tables(self.h, p(eng), p(cdf), p(tb))
        return eng, cdf, tb

    def emit_recomb(self, tree, ps, prev_rays, cut, nrec, tick, seed, boxsize, threshold=0.1, disable_lte=False,
                    lte=None):
        """the recombination rays of a tick from the previous tick's rays
        (or_emit_recomb); yGrec of ps is debited in place"""
        eng, cdf, tb = self.lte_tables() if lte is None else lte
        prev = np.ascontiguousarray(prev_rays, RAY_DTYPE)
        cap = max(1, abs(int(nrec)) * 3 if nrec > 0 else 3 * int(ps.npar))
        out = np.zeros(cap, RAY_DTYPE)
        s = psys_struct(ps)
        cut = min(int(cut), 2**62)
        n = self.L.or_emit_recomb(self.h, tree.h, C.byref(s), p(prev), len(prev), cut, int(nrec), float(threshold),
                                  int(bool(disable_lte)), p(eng), p(cdf), p(tb), float(boxsize), int(tick),
                                  int(seed) & 0xFFFFFFFF, p(out), cap)
        assert n >= 0
        return out[:n]

    # lookups
    def ar_get(self, col, logT):
        return self.L.or_ar_get(self.h, col, logT)

    def xs_get(self, i, e):
        return self.L.or_xs_get(self.h, i, e)

    def sph_depth(self, r):
        return self.L.or_sph_depth(self.h, r)

    def secion_get(self, i, E, lx):
        return self.L.or_secion_get(self.h, i, E, lx)

    def ar_table(self):
        n = self.L.or_ar_nrows(self.h)
        out = np.zeros(16 * n)
        mn, si = C.c_double(), C.c_double()
        self.L.or_ar_table(self.h, p(out), C.byref(mn), C.byref(si))
        return out.reshape(16, n), mn.value, si.value

    def ieye2T(self, ie, ye):
        return self.L.or_ieye2T(self.h, ie, ye)

    def Tye2ie(self, T, ye):
        return self.L.or_Tye2ie(self.h, T, ye)

    # chemistry
    def step_evolve(self, steps, rtol=1e-4, nthreads=8):
        s = np.ascontiguousarray(steps, STEP_DTYPE).copy()
        bad = self.L.or_step_evolve_many(self.h, p(s), len(s), rtol, nthreads)
        return s, bad

    def ros3_fixed(self, step, nsub):
        """(x[7], summed embedded error[7]) of nsub fixed ROS3 steps"""
        st = np.ascontiguousarray(step, STEP_DTYPE).reshape(1).copy()
        x = np.zeros(7)
        e = np.zeros(7)
        if self.L.or_ros3_fixed(self.h, p(st), nsub, p(x), p(e)):
            raise RuntimeError("singular ROS3 matrix")
        return x, e

    def chem_rhs(self, step, x):
        """or_chem_rhs of one step record at state x[7]"""
        st = np.ascontiguousarray(step, STEP_DTYPE).reshape(1).copy()
        data = (C.c_void_p * 2)(self.h, st.ctypes.data)
        xx = np.array(x, np.float64)
        f = np.zeros(7)
        self.L.or_chem_rhs(0.0, p(xx), p(f), C.cast(data, C.c_void_p))
        return f

    def chem_jac(self, step, x):
        """(f, J) of or_chem_jac"""
        st = np.ascontiguousarray(step, STEP_DTYPE).reshape(1).copy()
        xx = np.array(x, np.float64)
        f = np.zeros(7)
        J = np.zeros((7, 7))
        self.L.or_chem_jac(self.h, p(st), p(xx), p(f), p(J))
        return f, J

    def lsoda_evolve(self, steps):
        """the reference's LSODA (oracle/_ref) driven like step_evolve_lsoda
        with the oracle's right-hand side"""
        R = ref_lib()
        if R is None:
            raise RuntimeError("oracle/_ref is not built")
        out = np.ascontiguousarray(steps, STEP_DTYPE).copy()
        rhs = C.cast(self.L.or_chem_rhs, C.c_void_p)
        states = np.zeros(len(out), np.int32)
        for i in range(len(out)):
            st = out[i:i + 1]
            data = (C.c_void_p * 2)(self.h, st.ctypes.data)
            x = np.array([st["lambdaH"][0], st["lambdaHeI"][0], st["lambdaHeII"][0],
                          st["yGrecHII"][0], st["yGrecHeII"][0], st["yGrecHeIII"][0], st["ie"][0]])
            states[i] = R.ref_lsoda_evolve(rhs, C.cast(data, C.c_void_p), p(x), 1000)
            for k, name in enumerate(("lambdaH", "lambdaHeI", "lambdaHeII", "yGrecHII",
                                      "yGrecHeII", "yGrecHeIII", "ie")):
                out[name][i] = x[k]
        return out, states


class Tree:
    def __init__(self, pos, sml, boxsize):
        self.L = lib()
        self.pos = np.ascontiguousarray(pos, np.float32)
        self.sml = np.ascontiguousarray(sml, np.float32)
        self.h = self.L.or_tree_build(p(self.pos), p(self.sml), len(self.sml), boxsize)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_tree_free(self.h)
            self.h = None

    def export(self):
        n = self.L.or_tree_ncells(self.h)
        fc = np.zeros(n, np.int64)
        par = np.zeros(n, np.int64)
        npar = np.zeros(n, np.uint32)
        bot = np.zeros((n, 3), np.float32)
        top = np.zeros((n, 3), np.float32)
        dep = np.zeros(n, np.int32)
        self.L.or_tree_export(self.h, p(fc), p(par), p(npar), p(bot), p(top), p(dep))
        return dict(first_child=fc, parent=par, npar=npar, bot=bot, top=top, depth=dep)

    def leaf_of(self):
        out = np.zeros(len(self.sml), np.int64)
        self.L.or_tree_leaf_of(self.h, p(out))
        return out


def psys_struct(ps):
    """OrPsys view of a psphray_amd.ParticleSystem-like object (same fields)"""
    s = OrPsys()
    s.npar = ps.npar
    for name, _ in OrPsys._fields_[1:]:
        setattr(s, name, getattr(ps, name).ctypes.data)
    return s


def trace_rays(tree, ps, rays, nthreads=8):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    s = psys_struct(ps)
    off = np.zeros(len(rays) + 1, np.int64)
    L = lib()
    total = L.or_trace_rays(tree.h, C.byref(s), p(rays), len(rays), p(off), None, 0, nthreads)
    hits = np.zeros(max(total, 1), HIT_DTYPE)
    L.or_trace_rays(tree.h, C.byref(s), p(rays), len(rays), p(off), p(hits), total, nthreads)
    return off, hits[:total]


def deposit(orc, ps, rays, off, hits, nthreads=1):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    s = psys_struct(ps)
    st = OrStats()
    ln = np.zeros(len(rays))
    tm = np.zeros(len(rays))
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    lib().or_deposit(orc.h, C.byref(s), p(rays), len(rays), p(off), p(hits), p(ln), p(tm),
                     C.byref(st), nthreads)
    return ln, tm, st


def merge(ps, off, hits):
    s = OrPsys()
    s.npar = ps if isinstance(ps, (int, np.integer)) else ps.npar
    act = np.zeros(max(len(hits), 1), np.int64)
    n = lib().or_merge(C.byref(s), len(off) - 1, p(off), p(np.ascontiguousarray(hits, HIT_DTYPE)),
                       p(act))
    return act[:n]


def update(orc, ps, active, tick, tick_time, nthreads=8):
    s = psys_struct(ps)
    st = OrStats()
    active = np.ascontiguousarray(active, np.int64)
    rc = lib().or_update(orc.h, C.byref(s), p(active), len(active), tick, tick_time,
                         C.byref(st), nthreads)
    return rc, st


def tick(orc, tree, ps, rays, tick_no, tick_time, nthreads=8):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    s = psys_struct(ps)
    st = OrStats()
    rc = lib().or_tick(orc.h, tree.h, C.byref(s), p(rays), len(rays), tick_no, tick_time,
                       C.byref(st), nthreads)
    return rc, st


class Emitter:
    """the oracle's emit_rays over point sources (or_emit_rays): the same
    sources, spectra and lastemit bookkeeping as a RayEmitter"""

    def __init__(self, seed, pos, ngd, field, spec, spectra, boxsizes, tick_time):
        self.seed = int(seed) & 0xFFFFFFFF
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.ngd = np.ascontiguousarray(ngd, np.float64)
        self.field = np.ascontiguousarray(field, np.int32)
        self.spec = np.ascontiguousarray(spec, np.int32)
        self.off = np.zeros(len(spectra) + 1, np.int64)
        self.off[1:] = np.cumsum([len(e) for e, _ in spectra])
        self.eng = np.ascontiguousarray(np.concatenate([e for e, _ in spectra]), np.float64)
        self.cdf = np.ascontiguousarray(np.concatenate([c for _, c in spectra]), np.float64)
        self.boxsizes = np.ascontiguousarray(boxsizes, np.float64)
        self.tick_time = float(tick_time)
        self.lastemit = np.zeros(len(self.ngd), np.int64)

    def emit(self, tick, nray, per_field=None):
        pf = int(per_field or 0)
        cap = pf * len(self.boxsizes) if pf else int(nray)
        out = np.zeros(max(cap, 1), RAY_DTYPE)
        n = lib().or_emit_rays(len(self.ngd), p(self.pos), p(self.ngd), p(self.field), p(self.spec),
                               p(self.lastemit), p(self.off), p(self.eng), p(self.cdf), len(self.boxsizes),
                               p(self.boxsizes), int(tick), self.tick_time, int(nray), pf, self.seed,
                               p(out), cap)
        assert n >= 0
        return out[:n]
