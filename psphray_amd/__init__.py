"""psphray_amd -- MI355X (gfx950) photon-packet tracer for P-SPHRAY's hot path.

Python mirror of the C ABI in include/psphray_amd.h.  The heavy lifting is
in libpsphray_amd.so (hand-written HIP kernels + a C host layer); this module
only marshals numpy arrays.  Names follow the reference: a particle system
(PSystem, src/psystem.h:32-78) with its SoA fields, ray packets (struct r_t,
src/loop.c:21-35), ticks and epochs.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (PSR_ACCUM_ATOMIC, PSR_ACCUM_AUTO, PSR_ACCUM_BINNED, PSR_CHEM_EXACT,
                   PSR_CHEM_SURROGATE, PSR_CHEM_SURROGATE_STIFF, PSR_PHASE_ALL, PSR_PHASE_DEPOSIT, PSR_PHASE_TRACE,
                   PSR_PHASE_UPDATE, PSR_TUNE_DEPOSIT_ACCUM, PSR_TUNE_DEPOSIT_WALK,
                   PSR_TUNE_GRAPH_LAUNCHES, PSR_TUNE_TICK_GRAPH, PSR_TUNE_TICK_OVERLAP, PSR_TUNE_TRACE_THREADS,
                   PSR_WALK_AUTO, PSR_WALK_MULTI,
                   PSR_WALK_MULTI2, PSR_WALK_MULTI8, PSR_WALK_SINGLE, RAY_DTYPE, STEP_DTYPE, ptr)

__all__ = ["Physics", "Context", "ParticleSystem", "PsrError", "RAY_DTYPE", "STEP_DTYPE",
           "PSR_PHASE_ALL", "PSR_PHASE_TRACE", "PSR_PHASE_DEPOSIT", "PSR_PHASE_UPDATE",
           "PSR_CHEM_EXACT", "PSR_CHEM_SURROGATE", "PSR_CHEM_SURROGATE_STIFF", "PSR_TUNE_DEPOSIT_ACCUM", "PSR_ACCUM_AUTO",
           "PSR_ACCUM_ATOMIC", "PSR_ACCUM_BINNED", "PSR_TUNE_TICK_GRAPH", "PSR_TUNE_GRAPH_LAUNCHES",
           "PSR_TUNE_DEPOSIT_WALK", "PSR_TUNE_TRACE_THREADS", "PSR_TUNE_TICK_OVERLAP", "PSR_WALK_AUTO", "PSR_WALK_SINGLE", "PSR_WALK_MULTI", "PSR_WALK_MULTI2",
           "PSR_WALK_MULTI8",
           "default_units", "make_flags"]

# internal unit defaults, src/config.c:43-51
C_1_CMH = 3.0835455558318480e+21
C_1_GRAMH = 1.9847005219450705e+43
C_1_SECH = 3.08568025e+16


class PsrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"psphray_amd error {code}: {msg}")
        self.code = code


def default_units(h=0.72, hmf=0.76, hemf=None):
    """units_init (src/units.c:66-104); hemf defaults to 1 - hmf (src/config.c:113)"""
    u = _abi.Units()
    if hemf is None:
        hemf = 1.0 - hmf
    rc = _abi.lib().psr_units_init(C.byref(u), h, hmf, hemf, C_1_CMH, C_1_GRAMH, C_1_SECH)
    if rc:
        raise PsrError(rc, "psr_units_init")
    return u


def make_flags(isothermal=True, adiabatic=False, on_the_spot=True, h_only=True,
               secondary_ionization=True, trace_only=False, no_photon=False, comoving=False,
               redshift=0.0, fake_temperature=-1.0):
    return _abi.Flags(int(isothermal), int(adiabatic), int(on_the_spot), int(h_only),
                      int(secondary_ionization), int(trace_only), int(no_photon), int(comoving),
                      float(redshift), float(fake_temperature))


class Physics:
    """Units, run flags and host-built tables (rates file, cross sections,
    SPH column kernel, secondary ionization)."""

    def __init__(self, rates_path, units=None, flags=None, with_secion=False, kernel_path="default"):
        """kernel_path: the reference's SPH column table (src/kernel.akline
        format); "default" takes the copy beside the rates file when there is
        one, None keeps the quadrature table psr_tables_build computes"""
        self.units = units if units is not None else default_units()
        self.flags = flags if flags is not None else make_flags()
        self.tables = _abi.Tables()
        err = C.create_string_buffer(512)
        rc = _abi.lib().psr_tables_build(C.byref(self.tables), C.byref(self.units),
                                         str(rates_path).encode(), int(with_secion), err, 512)
        if rc:
            raise PsrError(rc, err.value.decode())
        if kernel_path == "default":
            cand = os.path.join(os.path.dirname(os.path.abspath(str(rates_path))), "kernel.akline")
            kernel_path = cand if os.path.exists(cand) else None
        if kernel_path is not None:
            rc = _abi.lib().psr_tables_load_kernel(C.byref(self.tables), str(kernel_path).encode(),
                                                   err, 512)
            if rc:
                raise PsrError(rc, err.value.decode())
        self.rates_path = str(rates_path)
        self.kernel_path = kernel_path

    def __del__(self):
        try:
            _abi.lib().psr_tables_free(C.byref(self.tables))
        except Exception:
            pass

    def ar_get(self, col, logT):
        return _abi.lib().psr_tab_ar_get(C.byref(self.tables), col, logT)

    def xs_get(self, i, eng):
        return _abi.lib().psr_tab_xs_get(C.byref(self.tables), i, eng)

    def sph_depth(self, r_h):
        return _abi.lib().psr_tab_sph_depth(C.byref(self.tables), r_h)

    def lte_tables(self):
        """LTE recombination spectra (psr_lte_tables): energies [3][1024],
        cdf [3][128][1024], row bounds in T [127]"""
        eng = np.zeros((3, _abi.PSR_LTE_NFREQ))
        cdf = np.zeros((3, _abi.PSR_LTE_NLOGT, _abi.PSR_LTE_NFREQ))
        tb = np.zeros(_abi.PSR_LTE_NLOGT - 1)
        rc = _abi.lib().psr_lte_tables(C.byref(self.tables), C.byref(self.units), eng.ctypes.data, cdf.ctypes.data,
                                       tb.ctypes.data)
        if rc:
            raise PsrError(rc, "psr_lte_tables")
        return eng, cdf, tb

    def kernel_table(self):
        n = self.tables.kernel_n
        return np.ctypeslib.as_array(self.tables.kernel, shape=(n,)).copy()

    def ieye2T(self, ie, ye):
        u = self.units
        return u.U_MPROTON_OVER_C_BOLTZMANN * ie / (u.C_HEMF * 0.25 + u.C_HMF + u.C_HMF * ye) * 2.0 / 3.0

    def Tye2ie(self, T, ye):
        u = self.units
        return T * u.C_BOLTZMANN_OVER_U_MPROTON * (u.C_HEMF * 0.25 + u.C_HMF + u.C_HMF * ye) / 2.0 * 3.0


_F32 = ("sml", "mass", "rho", "ie", "yeMET", "yGdepHI", "yGdepHeI", "yGdepHeII", "yGrecHII",
        "yGrecHeII", "yGrecHeIII", "heat")
_F64 = ("lambdaH", "lambdaHeI", "lambdaHeII")


class ParticleSystem:
    """SoA particle fields of PSystem (src/psystem.h:32-78) as numpy arrays."""

    def __init__(self, npar):
        self.npar = int(npar)
        self.pos = np.zeros((npar, 3), np.float32)
        for f in _F32:
            setattr(self, f, np.zeros(npar, np.float32))
        for f in _F64:
            setattr(self, f, np.zeros(npar, np.float64))
        self.flag = np.zeros(npar, np.int8)
        self.lasthit = np.zeros(npar, np.int64)
        self.hits = np.zeros(npar, np.int32)

    def copy(self):
        o = ParticleSystem(0)
        o.npar = self.npar
        for f in ("pos", "flag", "lasthit", "hits") + _F32 + _F64:
            setattr(o, f, getattr(self, f).copy())
        return o

    def fields(self):
        return ("pos", "flag", "lasthit", "hits") + _F32 + _F64

    def as_struct(self):
        s = _abi.State()
        s.npar = self.npar
        for f in ("pos",) + _F32:
            a = getattr(self, f)
            assert a.dtype == np.float32 and a.flags.c_contiguous, f
            setattr(s, f, a.ctypes.data_as(_abi._PF))
        for f in _F64:
            a = getattr(self, f)
            assert a.dtype == np.float64 and a.flags.c_contiguous, f
            setattr(s, f, a.ctypes.data_as(_abi._PD))
        s.flag = self.flag.ctypes.data_as(C.POINTER(C.c_int8))
        s.lasthit = self.lasthit.ctypes.data_as(C.POINTER(C.c_int64))
        s.hits = self.hits.ctypes.data_as(C.POINTER(C.c_int32))
        return s


class Context:
    """One device context (psr_ctx): tree, particle state, ray pool, stream."""

    def __init__(self, device=0):
        self._lib = _abi.lib()
        h = C.c_void_p()
        rc = self._lib.psr_create(C.byref(h), device)
        if rc:
            raise PsrError(rc, "psr_create failed (no usable GPU?)")
        self.h = h
        self.physics = None

    def close(self):
        if getattr(self, "h", None):
            self._lib.psr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc:
            msg = self._lib.psr_last_error(self.h)
            raise PsrError(rc, f"{what}: {msg.decode() if msg else ''}")

    def set_physics(self, physics):
        self.physics = physics
        self._check(self._lib.psr_set_physics(self.h, C.byref(physics.units),
                                              C.byref(physics.flags), C.byref(physics.tables)),
                    "psr_set_physics")

    def tree_build(self, pos, sml, flag=None, field_offsets=None, boxsizes=None, boxsize=None):
        pos = np.ascontiguousarray(pos, np.float32)
        sml = np.ascontiguousarray(sml, np.float32)
        n = len(sml)
        if field_offsets is None:
            field_offsets = np.array([0, n], np.int64)
            boxsizes = np.array([boxsize], np.float64)
        fo = np.ascontiguousarray(field_offsets, np.int64)
        bs = np.ascontiguousarray(boxsizes, np.float64)
        fl = None if flag is None else np.ascontiguousarray(flag, np.int8)
        self._check(self._lib.psr_tree_build(self.h, ptr(pos), ptr(sml), ptr(fl), n, len(bs),
                                             ptr(fo), ptr(bs)), "psr_tree_build")

    def tree_info(self):
        nn, dp, sk = C.c_int64(), C.c_int32(), C.c_int64()
        self._check(self._lib.psr_tree_info(self.h, C.byref(nn), C.byref(dp), C.byref(sk)),
                    "psr_tree_info")
        return nn.value, dp.value, sk.value

    def tree_export(self):
        nn, _, _ = self.tree_info()
        bot = np.zeros((nn, 3), np.float32)
        top = np.zeros((nn, 3), np.float32)
        fc = np.zeros(nn, np.int32)
        npar = np.zeros(nn, np.int32)
        lev = np.zeros(nn, np.int32)
        self._check(self._lib.psr_tree_export(self.h, nn, ptr(bot), ptr(top), ptr(fc), ptr(npar),
                                              ptr(lev)), "psr_tree_export")
        return dict(bot=bot, top=top, first_child=fc, npar=npar, level=lev)

    def load(self, ps, field_offsets=None, boxsizes=None, boxsize=None):
        """epoch switch: device tree build + state upload"""
        self.tree_build(ps.pos, ps.sml, ps.flag, field_offsets, boxsizes, boxsize)
        self.state_upload(ps)

    def state_upload(self, ps):
        s = ps.as_struct()
        self._check(self._lib.psr_state_upload(self.h, C.byref(s)), "psr_state_upload")

    def state_download(self, ps):
        s = ps.as_struct()
        self._check(self._lib.psr_state_download(self.h, C.byref(s)), "psr_state_download")
        return ps

    def rays_upload(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        self._check(self._lib.psr_rays_upload(self.h, ptr(rays), len(rays)), "psr_rays_upload")
        self._nray_pool = len(rays)

    def set_sources(self, emitter):
        """the sources and spectra of a RayEmitter (and its lastemit ticks)
        for device emission: psr_sources_upload + psr_spectra_upload"""
        pos = np.ascontiguousarray(emitter.pos, np.float64)
        ngd = np.ascontiguousarray(emitter.ngd, np.float64)
        fld = np.ascontiguousarray(emitter.field, np.int32)
        spc = np.ascontiguousarray(emitter.spec, np.int32)
        le = np.ascontiguousarray(emitter.lastemit, np.int64)
        self._check(self._lib.psr_sources_upload(self.h, len(ngd), ptr(pos), ptr(ngd), ptr(fld), ptr(spc),
                                                 ptr(le)), "psr_sources_upload")
        off = np.zeros(len(emitter.spectra) + 1, np.int64)
        off[1:] = np.cumsum([len(e) for e, _ in emitter.spectra])
        eng = np.ascontiguousarray(np.concatenate([e for e, _ in emitter.spectra]), np.float64)
        cdf = np.ascontiguousarray(np.concatenate([c for _, c in emitter.spectra]), np.float64)
        self._check(self._lib.psr_spectra_upload(self.h, len(emitter.spectra), ptr(off), ptr(eng), ptr(cdf)),
                    "psr_spectra_upload")
        self._emit_seed = emitter.seed
        self._emit_tick_time = emitter.tick_time

    def emit_rays(self, tick, nray, per_field=None, pool_first=0, seed=None, tick_time=None):
        """psr_emit_rays: the tick's source rays into the pool at pool_first;
        returns the ray count"""
        n = C.c_int64(0)
        seed = self._emit_seed if seed is None else seed
        tt = self._emit_tick_time if tick_time is None else tick_time
        self._check(self._lib.psr_emit_rays(self.h, int(tick), float(tt), int(nray), int(per_field or 0),
                                            int(seed) & 0xFFFFFFFF, int(pool_first), C.byref(n)),
                    "psr_emit_rays")
        self._nray_pool = max(getattr(self, "_nray_pool", 0), int(pool_first) + n.value)
        return n.value

    def output_gather(self, npar):
        """the output blocks the ticks change, gathered on the device
        (psr_output_gather): dict of xHI, xHeI, xHeII, ye, gammaHI, ie
        (float32), lasthit (int64), hits (int32)"""
        out = {k: np.zeros(npar, np.float32) for k in ("xHI", "xHeI", "xHeII", "ye", "gammaHI", "ie")}
        out["lasthit"] = np.zeros(npar, np.int64)
        out["hits"] = np.zeros(npar, np.int32)
        self._check(self._lib.psr_output_gather(self.h, *[out[k].ctypes.data for k in
                                                          ("xHI", "xHeI", "xHeII", "ye", "gammaHI", "ie",
                                                           "lasthit", "hits")]), "psr_output_gather")
        return out

    def tuning_save(self, path):
        """the decided launch choices into a file (psr_tuning_save)"""
        self._check(self._lib.psr_tuning_save(self.h, str(path).encode()), "psr_tuning_save")

    def tuning_load(self, path):
        """this workload's saved launch choices (psr_tuning_load)"""
        self._check(self._lib.psr_tuning_load(self.h, str(path).encode()), "psr_tuning_load")

    def set_recombination(self, nrec, threshold=0.1, disable_lte=False):
        """recombination rays appended by emit_rays (psr_set_recombination)"""
        self._check(self._lib.psr_set_recombination(self.h, int(nrec), float(threshold), int(bool(disable_lte))),
                    "psr_set_recombination")

    def sources_lastemit(self, nsrc):
        le = np.zeros(nsrc, np.int64)
        self._check(self._lib.psr_sources_lastemit(self.h, ptr(le)), "psr_sources_lastemit")
        return le

    def rays_download(self, first, n):
        """n rays of the pool from first (for tests): psr_rays_download"""
        out = np.zeros(n, RAY_DTYPE)
        self._check(self._lib.psr_rays_download(self.h, int(first), int(n), ptr(out)), "psr_rays_download")
        return out

    def trace_lists(self, ray_first=0, nray=None):
        """per-ray hit lists sorted by (d, ipar): offsets, ipar, d, b"""
        if nray is None:
            nray = self._nray_pool - ray_first
        off = np.zeros(nray + 1, np.int64)
        rc = self._lib.psr_trace_lists(self.h, ray_first, nray, ptr(off), 0, None, None, None)
        if rc not in (0, _abi.PSR_ERR_CAPACITY):
            self._check(rc, "psr_trace_lists")
        total = int(off[-1])
        ip = np.zeros(max(total, 1), np.int64)
        d = np.zeros(max(total, 1), np.float64)
        b = np.zeros(max(total, 1), np.float64)
        self._check(self._lib.psr_trace_lists(self.h, ray_first, nray, ptr(off), total, ptr(ip),
                                              ptr(d), ptr(b)), "psr_trace_lists")
        return off, ip[:total], d[:total], b[:total]

    def tick(self, tick, tick_time, ray_first=0, nray=None, phases=PSR_PHASE_ALL,
             chem=PSR_CHEM_EXACT):
        if nray is None:
            nray = self._nray_pool - ray_first
        self._check(self._lib.psr_tick(self.h, ray_first, nray, tick, tick_time, phases, chem),
                    "psr_tick")

    def ray_results(self, nray):
        ln = np.zeros(nray, np.float64)
        tm = np.zeros(nray, np.float64)
        self._check(self._lib.psr_ray_results(self.h, nray, ptr(ln), ptr(tm)), "psr_ray_results")
        return ln, tm

    def stats(self):
        st = _abi.Stats()
        self._check(self._lib.psr_get_stats(self.h, C.byref(st)), "psr_get_stats")
        return {k: getattr(st, k) for k, _ in _abi.Stats._fields_}

    def reset_stats(self):
        self._check(self._lib.psr_reset_stats(self.h), "psr_reset_stats")

    def set_timing(self, enable=True):
        self._check(self._lib.psr_set_timing(self.h, int(enable)), "psr_set_timing")

    def set_tuning(self, param, value):
        self._check(self._lib.psr_set_tuning(self.h, int(param), int(value)), "psr_set_tuning")

    def tuning(self, param=PSR_TUNE_DEPOSIT_ACCUM):
        """the choice in use (PSR_ACCUM_AUTO while the trial ticks run)"""
        v = C.c_int32(0)
        self._check(self._lib.psr_get_tuning(self.h, int(param), C.byref(v)), "psr_get_tuning")
        return v.value

    def timers(self):
        """accumulated ms: trace, sort, deposit, update, whole tick"""
        ms = np.zeros(5)
        self._check(self._lib.psr_get_timers(self.h, ptr(ms)), "psr_get_timers")
        return dict(zip(("trace", "sort", "deposit", "update", "tick"), ms.tolist()))

    def sync(self):
        self._check(self._lib.psr_sync(self.h), "psr_sync")

    # ---- learned chemistry (surrogate of step_evolve) ----
    def mlp_capture(self, cap):
        """exact chemistry of later ticks writes (features, targets) rows,
        up to cap of them, into a device buffer (0: off)"""
        self._check(self._lib.psr_mlp_capture(self.h, int(cap)), "psr_mlp_capture")

    def mlp_captured(self, max_rows=0):
        """(n, x[:max_rows], y[:max_rows]) of the captured rows"""
        n = C.c_int64(0)
        x = np.zeros((max(max_rows, 0), _abi.PSR_MLP_IN), np.float32)
        y = np.zeros((max(max_rows, 0), _abi.PSR_MLP_OUT), np.float32)
        self._check(self._lib.psr_mlp_captured(self.h, C.byref(n), ptr(x) if max_rows else None,
                                               ptr(y) if max_rows else None, int(max_rows)),
                    "psr_mlp_captured")
        k = min(n.value, max(max_rows, 0))
        return n.value, x[:k], y[:k]

    def mlp_train(self, x=None, y=None, n=0, epochs=1, batch=4096, lr=1e-3, seed=1):
        """train on (x, y) or, with x = y = None, on the captured rows;
        returns the last epoch's mean loss"""
        loss = C.c_float(0)
        if x is not None:
            x = np.ascontiguousarray(x, np.float32)
            y = np.ascontiguousarray(y, np.float32)
            n = len(x)
        self._check(self._lib.psr_mlp_train(self.h, ptr(x) if x is not None else None,
                                            ptr(y) if y is not None else None, int(n), int(epochs),
                                            int(batch), float(lr), int(seed), C.byref(loss)),
                    "psr_mlp_train")
        return loss.value

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), _abi.PSR_MLP_OUT), np.float32)
        self._check(self._lib.psr_mlp_forward(self.h, ptr(x), len(x), ptr(y)), "psr_mlp_forward")
        return y

    def mlp_params(self):
        p = np.zeros(self._lib.psr_mlp_param_count(), np.float32)
        self._check(self._lib.psr_mlp_get_params(self.h, ptr(p)), "psr_mlp_get_params")
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert len(p) == self._lib.psr_mlp_param_count()
        self._check(self._lib.psr_mlp_set_params(self.h, ptr(p)), "psr_mlp_set_params")

    def mlp_set_norm(self, mean, sigma, lo=None, hi=None):
        a = [np.ascontiguousarray(v, np.float32) if v is not None else None for v in (mean, sigma, lo, hi)]
        self._check(self._lib.psr_mlp_set_norm(self.h, *[ptr(v) if v is not None else None for v in a]),
                    "psr_mlp_set_norm")

    def mlp_norm(self):
        """(mean, sigma, lo, hi) of the inputs"""
        v = [np.zeros(_abi.PSR_MLP_IN, np.float32) for _ in range(4)]
        self._check(self._lib.psr_mlp_get_norm(self.h, *[ptr(a) for a in v]), "psr_mlp_get_norm")
        return tuple(v)

    def mlp_grad(self, x, y):
        """(gradient, loss) at the current parameters over all rows"""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self._lib.psr_mlp_param_count(), np.float32)
        loss = C.c_float(0)
        self._check(self._lib.psr_mlp_grad(self.h, ptr(x), ptr(y), len(x), ptr(g), C.byref(loss)),
                    "psr_mlp_grad")
        return g, loss.value

    def mlp_set_residual(self, tol):
        self._check(self._lib.psr_mlp_set_residual(self.h, float(tol)), "psr_mlp_set_residual")

    def mlp_set_balance_check(self, on=True):
        self._check(self._lib.psr_mlp_set_balance_check(self.h, int(bool(on))), "psr_mlp_set_balance_check")

    def solve_steps(self, steps, rtol=-1.0):
        steps = np.ascontiguousarray(steps, STEP_DTYPE).copy()
        self._check(self._lib.psr_solve_steps(self.h, ptr(steps), len(steps), rtol),
                    "psr_solve_steps")
        return steps
