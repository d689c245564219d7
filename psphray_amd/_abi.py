"""ctypes mirror of include/psphray_amd.h (the C ABI of libpsphray_amd.so)."""
import ctypes as C
import os

import numpy as np

PSR_OK = 0
PSR_ERR_ARG = 1
PSR_ERR_HIP = 2
PSR_ERR_CAPACITY = 3
PSR_ERR_TREE = 4
PSR_ERR_SOLVER = 5
PSR_ERR_STATE = 6
PSR_ERR_NOTSUP = 7

PSR_PHASE_TRACE = 1
PSR_PHASE_DEPOSIT = 2
PSR_PHASE_UPDATE = 4
PSR_PHASE_ALL = 7
PSR_CHEM_EXACT = 0
PSR_CHEM_SURROGATE = 1
PSR_CHEM_SURROGATE_STIFF = 2
PSR_MLP_IN = 22
PSR_NOALLOC_MAX_STEPS = 32768
PSR_LTE_NFREQ = 1024
PSR_LTE_NLOGT = 128
PSR_MLP_OUT = 4
PSR_TUNE_DEPOSIT_ACCUM = 1
PSR_ACCUM_AUTO = 0
PSR_ACCUM_ATOMIC = 1
PSR_ACCUM_BINNED = 2
PSR_TUNE_TICK_GRAPH = 2
PSR_TUNE_GRAPH_LAUNCHES = 3
PSR_TUNE_DEPOSIT_WALK = 4
PSR_TUNE_TRACE_THREADS = 5
PSR_TUNE_TICK_OVERLAP = 6
PSR_WALK_AUTO = 0
PSR_WALK_SINGLE = 1
PSR_WALK_MULTI = 2
PSR_WALK_MULTI2 = 3
PSR_WALK_MULTI8 = 4

_D = C.c_double
_I32 = C.c_int32
_I64 = C.c_int64
_PD = C.POINTER(C.c_double)
_PF = C.POINTER(C.c_float)


class Units(C.Structure):
    _fields_ = [(n, _D) for n in (
        "C_H C_HMF C_HEMF C_1_CMH C_1_GRAMH C_1_SECH "
        "U_CM U_GRAM U_SEC U_M U_KM U_KPC U_YR U_MYR U_KG U_MSUN U_MPROTON U_J U_EV U_ERG U_RY_ENG "
        "C_BOLTZMANN C_PLANCK C_SPEED_LIGHT C_H_PER_MASS C_HE_PER_MASS C_HI_ENERGY C_HEI_ENERGY "
        "C_HEII_ENERGY U_MPROTON_OVER_C_BOLTZMANN C_BOLTZMANN_OVER_U_MPROTON").split()]


class Flags(C.Structure):
    _fields_ = [(n, _I32) for n in
                "isothermal adiabatic on_the_spot h_only secondary_ionization trace_only "
                "no_photon comoving".split()] + [("redshift", _D), ("fake_temperature", _D)]


class Tables(C.Structure):
    _fields_ = [
        ("ar_nrows", _I32), ("ar_min", _D), ("ar_step_inv", _D), ("ar", _PD),
        ("xs_nrows", _I32), ("xs_min", _D), ("xs_step_inv", _D), ("xs_cut", _D * 3), ("xs", _PD),
        ("kernel_n", _I32), ("kernel", _PD),
        ("si_n1", _I32), ("si_n2", _I32), ("si_min1", _D), ("si_min2", _D),
        ("si_step1_inv", _D), ("si_step2_inv", _D),
        ("si_x", _PD), ("si_y", _PD), ("si_phi_hi", _PD), ("si_phi_hei", _PD), ("si_eh", _PD),
    ]


class State(C.Structure):
    _fields_ = [
        ("npar", _I64), ("pos", _PF), ("sml", _PF), ("mass", _PF), ("rho", _PF), ("ie", _PF),
        ("yeMET", _PF), ("flag", C.POINTER(C.c_int8)),
        ("lambdaH", _PD), ("lambdaHeI", _PD), ("lambdaHeII", _PD),
        ("yGdepHI", _PF), ("yGdepHeI", _PF), ("yGdepHeII", _PF),
        ("yGrecHII", _PF), ("yGrecHeII", _PF), ("yGrecHeIII", _PF),
        ("heat", _PF), ("lasthit", C.POINTER(_I64)), ("hits", C.POINTER(_I32)),
    ]


RAY_DTYPE = np.dtype([("s", "<f4", 3), ("dir", "<f4", 3), ("Nph", "<f8"), ("freq", "<f8"),
                      ("length", "<f8"), ("type", "<i4"), ("field", "<i4"),
                      ("src_or_par", "<i8")], align=True)
assert RAY_DTYPE.itemsize == 64

STEP_DTYPE = np.dtype([(n, "<f8") for n in
                       "lambdaH lambdaHeI lambdaHeII ie yeMET yGdepHI yGdepHeI yGdepHeII heat "
                       "time nH T yGrecHII yGrecHeII yGrecHeIII step_remain".split()] +
                      [("refined", "<i4"), ("nsteps", "<i4"), ("ipar", "<i8")], align=True)
assert STEP_DTYPE.itemsize == 144


class SnapHeader(C.Structure):
    """psr_snap_header: the psphray snapshot header record (256 bytes)"""
    _fields_ = [("N", C.c_uint32 * 6), ("mass", C.c_double * 6), ("time", C.c_double),
                ("redshift", C.c_double), ("flag_sfr", C.c_int32), ("flag_feedback", C.c_int32),
                ("Nparticle_total_low", C.c_uint32 * 6), ("flag_cool", C.c_int32), ("Nfiles", C.c_int32),
                ("boxsize", C.c_double), ("OmegaM", C.c_double), ("OmegaL", C.c_double), ("h", C.c_double),
                ("flag_sft", C.c_int32), ("flag_met", C.c_int32), ("Nparticle_total_high", C.c_uint32 * 6),
                ("flag_entropy", C.c_int32), ("OmegaB", C.c_double), ("nrays", C.c_uint64),
                ("flag_Hmf", C.c_int32), ("flag_Hemf", C.c_int32), ("flag_Helium", C.c_int32),
                ("flag_gammaHI", C.c_int32), ("flag_cloudy", C.c_int32), ("flag_eos", C.c_int32),
                ("unused", C.c_int32 * 4)]


class EpochSnap(C.Structure):
    _fields_ = [("npar", C.c_int64)] + [(n, C.c_void_p) for n in "id pos mass sml rho ie xHI ye".split()]


class EpochStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in "matched added lost removed".split()] + [("mean_shift", C.c_double)]


class Stats(C.Structure):
    _fields_ = [(n, _I64) for n in
                "ticks total_deposit_count saturated_deposit_count active_count solver_steps "
                "solver_failures surrogate_count surrogate_fallbacks deposit_processed".split()] + \
               [(n, _D) for n in
                "lost_photon_count_sum first_ionization_HI first_ionization_HeI "
                "first_ionization_HeII secondary_ionization_HI secondary_ionization_HeI".split()]


# every symbol the header declares, with its ctypes signature
_VP = C.c_void_p
SIGNATURES = {
    "psr_create": (C.c_int, [C.POINTER(_VP), C.c_int]),
    "psr_destroy": (None, [_VP]),
    "psr_last_error": (C.c_char_p, [_VP]),
    "psr_sync": (C.c_int, [_VP]),
    "psr_units_init": (C.c_int, [C.POINTER(Units), _D, _D, _D, _D, _D, _D]),
    "psr_tables_build": (C.c_int, [C.POINTER(Tables), C.POINTER(Units), C.c_char_p, C.c_int,
                                   C.c_char_p, C.c_size_t]),
    "psr_tables_free": (None, [C.POINTER(Tables)]),
    "psr_tables_load_kernel": (C.c_int, [C.POINTER(Tables), C.c_char_p, C.c_char_p, C.c_size_t]),
    "psr_tab_ar_get": (_D, [C.POINTER(Tables), C.c_int, _D]),
    "psr_tab_xs_get": (_D, [C.POINTER(Tables), C.c_int, _D]),
    "psr_tab_sph_depth": (_D, [C.POINTER(Tables), _D]),
    "psr_set_physics": (C.c_int, [_VP, C.POINTER(Units), C.POINTER(Flags), C.POINTER(Tables)]),
    "psr_tree_build": (C.c_int, [_VP, _VP, _VP, _VP, _I64, _I32, _VP, _VP]),
    "psr_tree_info": (C.c_int, [_VP, C.POINTER(_I64), C.POINTER(_I32), C.POINTER(_I64)]),
    "psr_tree_export": (C.c_int, [_VP, _I64, _VP, _VP, _VP, _VP, _VP]),
    "psr_state_upload": (C.c_int, [_VP, C.POINTER(State)]),
    "psr_state_download": (C.c_int, [_VP, C.POINTER(State)]),
    "psr_rays_upload": (C.c_int, [_VP, _VP, _I64]),
    "psr_trace_lists": (C.c_int, [_VP, _I64, _I64, _VP, _I64, _VP, _VP, _VP]),
    "psr_tick": (C.c_int, [_VP, _I64, _I64, _I64, _D, _I32, _I32]),
    "psr_ray_results": (C.c_int, [_VP, _I64, _VP, _VP]),
    "psr_get_stats": (C.c_int, [_VP, C.POINTER(Stats)]),
    "psr_reset_stats": (C.c_int, [_VP]),
    "psr_set_timing": (C.c_int, [_VP, C.c_int]),
    "psr_get_timers": (C.c_int, [_VP, _VP]),
    "psr_set_tuning": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "psr_get_tuning": (C.c_int, [_VP, C.c_int32, _VP]),
    "psr_solve_steps": (C.c_int, [_VP, _VP, _I64, _D]),
    "psr_solve_steps_noalloc": (C.c_int, [_VP, _VP, _I64, _D]),
    "psr_sources_upload": (C.c_int, [_VP, _I64, _VP, _VP, _VP, _VP, _VP]),
    "psr_spectra_upload": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    "psr_emit_rays": (C.c_int, [_VP, _I64, _D, _I64, _I64, C.c_uint32, _I64, C.POINTER(C.c_int64)]),
    "psr_sources_lastemit": (C.c_int, [_VP, _VP]),
    "psr_rays_download": (C.c_int, [_VP, _I64, _I64, _VP]),
    "psr_set_recombination": (C.c_int, [_VP, _I64, _D, C.c_int32]),
    "psr_output_gather": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "psr_tuning_save": (C.c_int, [_VP, C.c_char_p]),
    "psr_tuning_load": (C.c_int, [_VP, C.c_char_p]),
    "psr_lte_tables": (C.c_int, [C.POINTER(Tables), C.POINTER(Units), _VP, _VP, _VP]),
    "psr_snapshot_block_info": (C.c_int, [C.c_char_p, C.POINTER(SnapHeader), C.c_char_p, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "psr_snapshot_read_header": (C.c_int, [C.c_char_p, C.POINTER(SnapHeader), C.c_char_p, C.c_size_t]),
    "psr_snapshot_read_block": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(SnapHeader), C.c_char_p, _VP,
                                          C.c_char_p, C.c_size_t]),
    "psr_snapshot_write": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(SnapHeader), C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_void_p), C.c_int32, C.c_char_p, C.c_size_t]),
    "psr_epoch_load": (C.c_int, [C.POINTER(State), _VP, _I64, C.POINTER(EpochSnap), C.POINTER(Units),
                                 C.c_int32, C.c_char_p, C.c_size_t]),
    "psr_epoch_switch": (C.c_int, [C.POINTER(State), _VP, _I64, C.POINTER(EpochSnap), C.POINTER(Units),
                                   C.c_int32, C.c_int32, _D, _D, _I64, _D, C.POINTER(EpochStats),
                                   C.c_char_p, C.c_size_t]),
    "psr_mlp_features": (C.c_int, [_VP, C.POINTER(Units), _VP]),
    "psr_mlp_train": (C.c_int, [_VP, _VP, _VP, _I64, _I32, _I32, C.c_float, C.c_uint64, _VP]),
    "psr_mlp_forward": (C.c_int, [_VP, _VP, _I64, _VP]),
    "psr_mlp_param_count": (_I64, []),
    "psr_mlp_get_params": (C.c_int, [_VP, _VP]),
    "psr_mlp_set_params": (C.c_int, [_VP, _VP]),
    "psr_mlp_grad": (C.c_int, [_VP, _VP, _VP, _I64, _VP, _VP]),
    "psr_mlp_set_residual": (C.c_int, [_VP, _D]),
    "psr_mlp_set_balance_check": (C.c_int, [_VP, _I32]),
    "psr_mlp_capture": (C.c_int, [_VP, _I64]),
    "psr_mlp_captured": (C.c_int, [_VP, _VP, _VP, _VP, _I64]),
    "psr_mlp_set_norm": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "psr_mlp_get_norm": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpsphray_amd.so")
_lib = None


def lib():
    """Load the in-tree HIP library; fails loudly when it is not built."""
    global _lib
    if _lib is None:
        # PSR_LIBRARY: another build of the same ABI (profiling builds)
        path = os.environ.get("PSR_LIBRARY") or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with psphray_amd/build.sh "
                               "(or __graft_entry__.build())")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """data pointer of a numpy array (or None)"""
    return None if a is None else C.c_void_p(a.ctypes.data)
