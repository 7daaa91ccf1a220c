"""ctypes binding of the C ABI declared in include/ffs/ffs.h.

The product path has no fallback: if libffs_hip.so is missing or fails to
load, every call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FFS_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FFS_LIB", os.path.join(HERE, "libffs_hip.so"))

FFS_OK = 0
ERRORS = {
    -1: "FFS_ERR_ARG",
    -2: "FFS_ERR_MEMORY",
    -3: "FFS_ERR_IO",
    -4: "FFS_ERR_NONFINITE",
    -5: "FFS_ERR_HIP",
    -6: "FFS_ERR_STATE",
}
FFS_AB3, FFS_ETD3, FFS_EULER = 0, 1, 2
FFS_FP64, FFS_FP32 = 0, 1
FFS_LES_NONE, FFS_LES_SMAGORINSKY = 0, 1
FFS_LES_VREMAN, FFS_LES_SILVIS, FFS_LES_PRODUCTION, FFS_LES_STABLE_NL = 2, 3, 4, 5
FFS_LES_FAKE_SMAGORINSKY, FFS_LES_SANDP, FFS_LES_DYNAMIC_SMAGORINSKY, FFS_LES_PIOMELLI_SMAGORINSKY = 6, 7, 8, 9
FFS_LES_DYNAMIC_SANDP, FFS_LES_PIOMELLI_SANDP = 10, 11
# lesModel / sLesModel values of the `global` file (les_types.jl:364-404)
LES_MODELS = {"Smagorinsky": FFS_LES_SMAGORINSKY, "Vreman": FFS_LES_VREMAN, "Silvis": FFS_LES_SILVIS,
              "productionViscosity": FFS_LES_PRODUCTION, "stableNonLinear": FFS_LES_STABLE_NL,
              "fakeSmagorinsky": FFS_LES_FAKE_SMAGORINSKY, "SandP": FFS_LES_SANDP,
              "dynamicSmagorinsky": FFS_LES_DYNAMIC_SMAGORINSKY,
              "piomelliSmagorinsky": FFS_LES_PIOMELLI_SMAGORINSKY, "dynamicSandP": FFS_LES_DYNAMIC_SANDP,
              "piomelliSandP": FFS_LES_PIOMELLI_SANDP}
FFS_FORCING_NONE, FFS_FORCING_SHELL, FFS_FORCING_RF, FFS_FORCING_QF = 0, 1, 2, 3
FFS_FORCING_CRF, FFS_FORCING_NRF, FFS_FORCING_CNRF, FFS_FORCING_A = 4, 5, 6, 7
FFS_HV_NONE, FFS_HV_HYPER, FFS_HV_BARRIER = 0, 1, 2
FFS_EQ_ROTATIONAL, FFS_EQ_NS = 0, 1
FFS_FSTATE_R, FFS_FSTATE_EM, FFS_FSTATE_KH, FFS_FSTATE_EF, FFS_FSTATE_N = 0, 1, 2, 3, 4
FFS_FIELD_U1, FFS_FIELD_U2, FFS_FIELD_U3, FFS_FIELD_RHO, FFS_FIELD_SCALAR = 0, 1, 2, 3, 4
FFS_FIELD_T11, FFS_FIELD_T12, FFS_FIELD_T13, FFS_FIELD_T22, FFS_FIELD_T23 = 5, 6, 7, 8, 9
FFS_FIELD_DFLUX1, FFS_FIELD_DFLUX2, FFS_FIELD_DFLUX3, FFS_FIELD_CS, FFS_FIELD_CP = 10, 11, 12, 13, 14
FFS_SPACE_REAL, FFS_SPACE_FOURIER, FFS_SPACE_FOURIER_TO_REAL = 0, 1, 2
FFS_SCALAR_ITERATION, FFS_SCALAR_TIME, FFS_SCALAR_DT, FFS_SCALAR_UMAX = 0, 1, 2, 3
FFS_TABLE_STATS, FFS_TABLE_SCALES, FFS_TABLE_NUMBERS = 0, 1, 2
FFS_EVENT_NONE, FFS_EVENT_OUTPUT = 0, 1
# the spectra quantities of a spectra record, in record order
SPEC_QUANTITIES = ("u_h", "u_v", "vis_h", "vis_v", "nl_h", "nl_v", "press_h", "press_v", "les_h", "les_v",
                   "buoyancy_v", "rho", "rhodiss", "rhonl", "forcing_h", "hvis_h", "hvis_v")
FFS_SPEC_Q = len(SPEC_QUANTITIES)


class FfsParams(ctypes.Structure):
    """Mirror of `ffs_params` (include/ffs/ffs.h)."""

    _fields_ = [
        ("size", ctypes.c_uint32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("x_domain", ctypes.c_double),
        ("y_domain", ctypes.c_double),
        ("z_domain", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("integrator", ctypes.c_int32),
        ("variable_dt", ctypes.c_int32),
        ("cfl", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("start", ctypes.c_int64),
        ("start_time", ctypes.c_double),
        ("dt_stat", ctypes.c_int32),
        ("dt_spec", ctypes.c_int32),
        ("write_time", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("les_model", ctypes.c_int32),
        ("smagorinsky_c", ctypes.c_double),
        ("filter_width", ctypes.c_double),
        ("density", ctypes.c_int32),
        ("pr", ctypes.c_double),
        ("density_gradient", ctypes.c_double),
        ("z_acceleration", ctypes.c_double),
        ("reference_density", ctypes.c_double),
        ("forcing", ctypes.c_int32),
        ("forcing_kf", ctypes.c_double),
        ("forcing_power", ctypes.c_double),
        ("forcing_tf", ctypes.c_double),
        ("forcing_alpha", ctypes.c_double),
        ("forcing_zf_tanh", ctypes.c_int32),
        ("alvelius_p", ctypes.c_double),
        ("alvelius_c", ctypes.c_double),
        ("alvelius_kf", ctypes.c_double),
        ("hyperviscosity", ctypes.c_int32),
        ("hv_nu", ctypes.c_double),
        ("hv_m", ctypes.c_int32),
        ("sb_initk", ctypes.c_double),
        ("sb_endk", ctypes.c_double),
        ("passive_scalar", ctypes.c_int32),
        ("scalar_pr", ctypes.c_double),
        ("scalar_gradient", ctypes.c_double),
        ("scalar_direction", ctypes.c_int32),
        ("equation", ctypes.c_int32),
        ("les_cp", ctypes.c_double),
        ("les_s_model", ctypes.c_int32),
        ("les_test_width", ctypes.c_double),
        ("les_cmin", ctypes.c_double),
        ("scalar_les", ctypes.c_int32),
        ("scalar_les_c", ctypes.c_double),
        ("density_les", ctypes.c_int32),
        ("density_les_c", ctypes.c_double),
        ("gravity_direction", ctypes.c_int32),
        ("density_hv", ctypes.c_int32),
        ("density_hv_nu", ctypes.c_double),
        ("density_hv_m", ctypes.c_int32),
        ("density_sb_initk", ctypes.c_double),
        ("density_sb_endk", ctypes.c_double),
    ]


class FfsPlan(ctypes.Structure):
    """ffs_plan of include/ffs/ffs.h."""

    _fields_ = [("x_pair", ctypes.c_int), ("x_waves", ctypes.c_int), ("x_split", ctypes.c_int),
                ("fusion", ctypes.c_int), ("z_slab", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


# name -> (restype, argtypes); every symbol declared in include/ffs/ffs.h
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "ffs_abi_version": (ctypes.c_int, []),
    "ffs_params_default": (ctypes.c_int, [ctypes.POINTER(FfsParams)]),
    "ffs_create": (ctypes.c_int, [ctypes.POINTER(FfsParams), ctypes.POINTER(_P)]),
    "ffs_set_field": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, ctypes.c_size_t]),
    "ffs_get_field": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, ctypes.c_size_t]),
    "ffs_get_field_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_size_t]),
    "ffs_seed_velocity": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    "ffs_seed_density": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    "ffs_initialize": (ctypes.c_int, [_P]),
    "ffs_advance": (ctypes.c_int, [_P, ctypes.c_int64, _I64, _I]),
    "ffs_writestats_now": (ctypes.c_int, [_P]),
    "ffs_checkpoint_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "ffs_forcing_state": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _I]),
    "ffs_checkpoint_save": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "ffs_checkpoint_load": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t]),
    "ffs_stats_pop": (ctypes.c_int, [_P, _D, ctypes.c_int, _I, _D, ctypes.c_int, _I, _D, ctypes.c_int, _I]),
    "ffs_stats_header": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "ffs_spectrum": (ctypes.c_int, [_P, _D, ctypes.c_int, _I, _D, ctypes.c_int, _I]),
    "ffs_spectra_axes": (ctypes.c_int, [_P, _D, ctypes.c_int, _I, _D, ctypes.c_int, _I]),
    "ffs_spectra_pop": (ctypes.c_int, [_P, _I64, _D, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "ffs_get_scalar": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "ffs_profile_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "ffs_profile_get": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_char_p), _D, _I64, ctypes.c_int, _I]),
    "ffs_profile_reset": (ctypes.c_int, [_P]),
    "ffs_device_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "ffs_synchronize": (ctypes.c_int, [_P]),
    "ffs_footprint": (ctypes.c_int, [ctypes.POINTER(FfsParams), ctypes.POINTER(ctypes.c_size_t)]),
    "ffs_device_memory": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "ffs_set_graph": (ctypes.c_int, [_P, ctypes.c_int]),
    "ffs_set_fusion": (ctypes.c_int, [_P, ctypes.c_int]),
    "ffs_get_plan": (ctypes.c_int, [_P, ctypes.POINTER(FfsPlan)]),
    "ffs_set_plan": (ctypes.c_int, [_P, ctypes.POINTER(FfsPlan)]),
    "ffs_plan_options": (ctypes.c_int, [_P, ctypes.POINTER(FfsPlan)]),
    "ffs_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "ffs_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
    "ffs_batch_advance": (ctypes.c_int, [_P, ctypes.c_int64, _I64, _I, _I]),
    "ffs_batch_synchronize": (ctypes.c_int, [_P]),
    "ffs_batch_set_merge": (ctypes.c_int, [_P, ctypes.c_int]),
    "ffs_batch_query": (ctypes.c_int, [_P]),
    "ffs_batch_last_advance_ms": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), _I64]),
    "ffs_batch_launch_info": (ctypes.c_int, [_P, _I, _I, _I]),
    "ffs_batch_last_error": (ctypes.c_char_p, [_P]),
    "ffs_batch_destroy": (None, [_P]),
    "ffs_last_error": (ctypes.c_char_p, [_P]),
    "ffs_destroy": (None, [_P]),
}

_lib = None


def load():
    """Loads libffs_hip.so once; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libffs_hip.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FfsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(rc, handle=None):
    if rc is not None and rc < 0:
        lib = load()
        msg = lib.ffs_last_error(handle)
        raise FfsError(rc, msg.decode() if msg else "")
    return rc


def default_params():
    p = FfsParams()
    check(load().ffs_params_default(ctypes.byref(p)))
    return p
