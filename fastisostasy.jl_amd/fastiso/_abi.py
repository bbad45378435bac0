"""ctypes declaration of the C ABI in include/fastiso.h and the loader of the HIP engine.

The engine library is fastisostasy.jl_amd/lib/libfastiso_hip.so, built in-tree by
fastisostasy.jl_amd/build.py. There is no fallback: if the library is missing or fails to load,
load_engine() raises.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE_PATH = os.path.join(PKG_ROOT, "lib", "libfastiso_hip.so")

FI_ABI_VERSION = 6

# status codes
FI_OK = 0
FI_ERR_INVALID = -1
FI_ERR_UNSUPPORTED = -2
FI_ERR_TIME_RANGE = -3
FI_ERR_NONFINITE = -4
FI_ERR_DEVICE = -5
FI_ERR_STATE = -6
FI_ERR_DT_MIN = -7

# model enums
FI_RIGID_LITHOSPHERE, FI_LATERALLY_CONSTANT_LITHOSPHERE, FI_LATERALLY_VARIABLE_LITHOSPHERE = 0, 1, 2
FI_RIGID_MANTLE, FI_RELAXED_MANTLE, FI_MAXWELL_MANTLE = 0, 1, 2
FI_INCOMPRESSIBLE_LITHOSPHERE_COLUMN, FI_COMPRESSIBLE_LITHOSPHERE_COLUMN = 0, 1
FI_LATERALLY_CONSTANT_SEA_SURFACE, FI_LATERALLY_VARIABLE_SEA_SURFACE, FI_IMPOSED_SEA_SURFACE = 0, 1, 2
FI_NO_SEALEVEL_LOAD, FI_INTERACTIVE_SEALEVEL_LOAD = 0, 1
FI_NO_CONTRIBUTION, FI_GOELZER_CONTRIBUTION = 0, 1

# field ids
FIELDS = {
    "litho_rigidity": 1, "effective_viscosity": 2, "scaled_pseudodiff_inv": 3, "maskactive": 4,
    "green_elastic": 5, "green_geoid": 6, "ref_u": 7, "ref_ue": 8, "ref_dz_ss": 9, "ref_z_b": 10,
    "ref_H_ice": 11, "distortion": 12, "green_viscous": 13, "tau": 14, "litho_thickness": 15,
    "u": 32, "dudt": 33, "ue": 34, "dz_ss": 35, "z_ss": 36, "z_b": 37, "H_ice": 38, "H_af": 39,
    "H_water": 40, "maskgrounded": 41, "maskocean": 42, "canom_ice": 43, "canom_seawater": 44,
    "canom_litho": 45, "canom_mantle": 46, "canom_load": 47, "canom_full": 48, "u_eq": 49, "u_x": 50, "u_y": 51,
}


class FiConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("Wx", C.c_double), ("Wy", C.c_double),
        ("g", C.c_double), ("seconds_per_year", C.c_double), ("rho_ice", C.c_double),
        ("rho_water", C.c_double), ("rho_seawater", C.c_double), ("r_pole", C.c_double),
        ("mE", C.c_double),
        ("litho_poissonratio", C.c_double), ("rho_uppermantle", C.c_double), ("rho_litho", C.c_double),
        ("lithosphere", C.c_int32), ("mantle", C.c_int32), ("lithosphere_column", C.c_int32),
        ("sea_surface", C.c_int32), ("sealevel_load", C.c_int32), ("volume_contribution", C.c_int32),
        ("density_contribution", C.c_int32), ("adjustment_contribution", C.c_int32),
        ("bsl_z", C.c_double),
        ("viscous_x_border", C.c_double), ("elastic_x_border", C.c_double), ("ssh_x_border", C.c_double),
        ("dt_sparse", C.c_double), ("reltol", C.c_double), ("abstol", C.c_double),
        ("device", C.c_int32), ("reserved", C.c_int32), ("dt_min", C.c_double),
    ]


class FiStats(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double),
        ("n_rhs", C.c_int64), ("n_accept", C.c_int64), ("n_reject", C.c_int64),
        ("count_sparse_updates", C.c_int64), ("n_output", C.c_int64),
        ("V_af", C.c_double), ("V_den", C.c_double), ("V_pov", C.c_double), ("delta_V", C.c_double),
        ("z_bsl", C.c_double), ("last_EEst", C.c_double),
        ("device_ms", C.c_double), ("n_launch", C.c_int64), ("n_accept_timed", C.c_int64),
        ("A_ocean", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FiBatchInfo(C.Structure):
    _fields_ = [
        ("n_members", C.c_int32), ("team_size", C.c_int32), ("n_teams", C.c_int32), ("tuned", C.c_int32),
        ("device_ms", C.c_double), ("n_launch", C.c_int64), ("cell_attempts", C.c_double),
        ("last_team_size", C.c_int32), ("n_candidates", C.c_int32),
        ("cand_size", C.c_int32 * 8), ("cand_rate", C.c_double * 8),
        ("wg_threads", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        n = self.n_candidates
        d["cand_size"] = list(self.cand_size)[:n]
        d["cand_rate"] = list(self.cand_rate)[:n]
        return d


class FiRkcInfo(C.Structure):
    _fields_ = [("rho", C.c_double), ("h", C.c_double), ("substeps", C.c_int32), ("stages", C.c_int32),
                ("n_rhs", C.c_int64), ("graph", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


FI_TUNE_WORKGROUPS, FI_TUNE_FFT_PLANS = 1, 2


class FiTuneInfo(C.Structure):
    _fields_ = [("workgroups", C.c_int32), ("n_wg_candidates", C.c_int32), ("wg_candidates", C.c_int32 * 8),
                ("wg_us", C.c_double * 8), ("plans", C.c_uint64 * 4), ("n_plan_candidates", C.c_int32 * 4),
                ("plan_default_us", C.c_double * 4), ("plan_best_us", C.c_double * 4)]

    def as_dict(self):
        n = self.n_wg_candidates
        return {"workgroups": self.workgroups, "wg_candidates": list(self.wg_candidates)[:n],
                "wg_us": list(self.wg_us)[:n], "plans": list(self.plans),
                "plan_radices": [plan_radices(c) for c in self.plans],
                "n_plan_candidates": list(self.n_plan_candidates), "plan_default_us": list(self.plan_default_us),
                "plan_best_us": list(self.plan_best_us)}


def plan_radices(code):
    """The radix of each pass of an FFT plan code (4 bits per pass, first pass lowest)."""
    out = []
    while code:
        out.append(int(code & 15))
        code >>= 4
    return out


def plan_code(radices):
    code = 0
    for s, r in enumerate(radices):
        code |= int(r) << (4 * s)
    return code


OUTPUT_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_int32)

_dp = C.POINTER(C.c_double)


def declare(lib):
    """Attach argtypes/restype of every fi_* entry point to a loaded library."""
    P = C.c_void_p
    sig = {
        "fi_abi_version": (C.c_int32, []),
        "fi_engine_name": (C.c_char_p, []),
        "fi_default_config": (None, [C.POINTER(FiConfig), C.c_int32, C.c_int32, C.c_double, C.c_double]),
        "fi_create": (C.c_int, [C.POINTER(FiConfig), C.POINTER(P)]),
        "fi_destroy": (None, [P]),
        "fi_last_error": (C.c_char_p, [P]),
        "fi_set_field": (C.c_int, [P, C.c_int32, _dp, C.c_size_t]),
        "fi_get_field": (C.c_int, [P, C.c_int32, _dp, C.c_size_t]),
        "fi_set_field_dev": (C.c_int, [P, C.c_int32, C.c_void_p, C.c_size_t]),
        "fi_get_field_dev": (C.c_int, [P, C.c_int32, C.c_void_p, C.c_size_t]),
        "fi_set_ice_snapshots": (C.c_int, [P, _dp, C.c_int32, _dp, C.c_int32]),
        "fi_set_sea_surface_snapshots": (C.c_int, [P, _dp, C.c_int32, _dp, C.c_int32]),
        "fi_set_timer_history": (C.c_int, [P, C.c_int64]),
        "fi_get_timer_history": (C.c_int, [P, _dp, _dp, C.c_int64, C.POINTER(C.c_int64)]),
        "fi_set_ice_separable": (C.c_int, [P, _dp, _dp, C.c_int32, _dp, C.c_int32]),
        "fi_set_bsl": (C.c_int, [P, C.c_int32, C.c_int32, C.c_double, _dp, _dp, C.c_int32, C.c_int32]),
        "fi_set_output_times": (C.c_int, [P, _dp, C.c_int32]),
        "fi_init": (C.c_int, [P, C.c_double, C.c_double]),
        "fi_rhs": (C.c_int, [P, C.c_double, _dp, _dp]),
        "fi_rhs_device": (C.c_int, [P, C.c_double, C.c_void_p, C.c_void_p]),
        "fi_sparse_update": (C.c_int, [P, C.c_double]),
        "fi_integrator_init": (C.c_int, [P]),
        "fi_integrate": (C.c_int, [P, C.c_double, OUTPUT_CB, C.c_void_p]),
        "fi_get_stats": (C.c_int, [P, C.POINTER(FiStats)]),
        "fi_set_time": (C.c_int, [P, C.c_double]),
        "fi_integrate_rkc": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.POINTER(FiRkcInfo)]),
        "fi_autotune": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(FiTuneInfo)]),
        "fi_set_fft_plans": (C.c_int, [P, C.POINTER(C.c_uint64)]),
        "fi_get_fft_plans": (C.c_int, [P, C.POINTER(C.c_uint64)]),
        "fi_set_workgroups": (C.c_int, [P, C.c_int32]),
        "fi_batch_autotune_fft": (C.c_int, [P, C.c_int32]),
        "fi_checkpoint_size": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "fi_checkpoint_save": (C.c_int, [P, C.c_void_p, C.c_size_t]),
        "fi_checkpoint_load": (C.c_int, [P, C.c_void_p, C.c_size_t]),
        "fi_create_batch": (C.c_int, [C.POINTER(FiConfig), C.c_int32, C.POINTER(P)]),
        "fi_batch_member": (P, [P, C.c_int32]),
        "fi_batch_destroy": (None, [P]),
        "fi_batch_last_error": (C.c_char_p, [P]),
        "fi_batch_set_team_size": (C.c_int, [P, C.c_int32]),
        "fi_batch_integrate": (C.c_int, [P, C.c_double, C.c_int32]),
        "fi_batch_integrate_each": (C.c_int, [P, _dp, C.c_int32]),
        "fi_batch_integrate_each_noalloc": (C.c_int, [P, _dp, C.c_int32]),
        "fi_batch_get_info": (C.c_int, [P, C.POINTER(FiBatchInfo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class FnoConfig(C.Structure):
    _fields_ = [
        ("in_channels", C.c_int32), ("width", C.c_int32), ("modes_x", C.c_int32), ("modes_y", C.c_int32),
        ("layers", C.c_int32), ("proj_hidden", C.c_int32), ("device", C.c_int32), ("reserved", C.c_int32),
        ("seed", C.c_uint64),
    ]


def declare_surrogate(lib):
    """argtypes/restype of the fi_surrogate_* entry points (include/fastiso_surrogate.h)."""
    P = C.c_void_p
    _fp = C.POINTER(C.c_float)
    sig = {
        "fi_fno_default_config": (None, [C.POINTER(FnoConfig)]),
        "fi_surrogate_create": (C.c_int, [C.POINTER(FnoConfig), C.POINTER(P)]),
        "fi_surrogate_destroy": (None, [P]),
        "fi_surrogate_last_error": (C.c_char_p, [P]),
        "fi_surrogate_num_params": (C.c_int64, [P]),
        "fi_surrogate_get_params": (C.c_int, [P, _fp, C.c_int64]),
        "fi_surrogate_set_params": (C.c_int, [P, _fp, C.c_int64]),
        "fi_surrogate_set_grid": (C.c_int, [P, C.c_int32, C.c_int32]),
        "fi_surrogate_forward": (C.c_int, [P, _fp, _fp]),
        "fi_surrogate_loss_grad": (C.c_int, [P, _fp, _fp, _dp, _fp]),
        "fi_surrogate_loss_grad_n": (C.c_int, [P, C.c_int32, _fp, _fp, _dp, _fp]),
        "fi_surrogate_set_data": (C.c_int, [P, C.c_int32, _fp, _fp]),
        "fi_surrogate_train": (C.c_int, [P, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_uint64, _dp]),
        "fi_surrogate_device_ms": (C.c_double, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_ENGINE = None


ENGINE_NAME = b"hip-gfx950"


def load_engine(path=None, allow_other_engine=False):
    """Load the HIP engine (libfastiso_hip.so). Raises if it is missing, if its ABI version differs,
    or if the library is another implementation of the ABI (the CPU oracle exports the same
    symbols): there is no CPU fallback. allow_other_engine lifts the last check for callers that
    load another implementation on purpose."""
    global _ENGINE
    if path is None and _ENGINE is not None:
        return _ENGINE
    # FASTISO_HIP_LIB selects another build of the same library (e.g. an instrumented one)
    p = path or os.environ.get("FASTISO_HIP_LIB") or ENGINE_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP engine not built: {p} is missing. Build it with "
            f"`python fastisostasy.jl_amd/build.py` (hipcc --offload-arch=gfx950).")
    lib = C.CDLL(p)
    for sym in ("fi_abi_version", "fi_engine_name"):
        if not hasattr(lib, sym):
            raise RuntimeError(f"{p} does not export {sym}: not a build of include/fastiso.h")
    lib.fi_abi_version.restype = C.c_int32
    lib.fi_engine_name.restype = C.c_char_p
    if lib.fi_abi_version() != FI_ABI_VERSION:
        raise RuntimeError(f"{p}: ABI version {lib.fi_abi_version()}, expected {FI_ABI_VERSION}")
    name = lib.fi_engine_name()
    if name != ENGINE_NAME and not allow_other_engine:
        raise RuntimeError(f"{p} is the {name.decode()} engine, not {ENGINE_NAME.decode()}")
    lib = declare(lib)
    if name == ENGINE_NAME:
        declare_surrogate(lib)
    lib._fastiso_path = os.path.abspath(p)
    if path is None:
        _ENGINE = lib
    return lib


def engine_identity(lib):
    """(engine name, absolute path of the loaded library) of a library returned by load_engine."""
    return lib.fi_engine_name().decode(), getattr(lib, "_fastiso_path", None)
