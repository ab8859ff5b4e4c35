"""ctypes binding of libshakti_hip.so (include/shakti_hip.h).

The library is loaded from this package's directory.  There is no fallback:
if the shared object is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libshakti_hip.so"

SHK_OK = 0
ERRORS = {-1: "SHK_ERR_ARG", -2: "SHK_ERR_HIP", -3: "SHK_ERR_NUMERIC", -4: "SHK_ERR_STATE", -5: "SHK_ERR_NODEVICE"}

FIELD_NAMES = ["h", "pw", "po", "b", "beta", "abs_ub", "mdot", "shear", "potential", "sensible", "Re", "K", "G", "zb",
               "zs", "H", "ieb", "lambda", "A_visc", "N", "mask",
               "dhdx", "q_x", "Re_x", "b_x", "ub_x", "taub_x", "dpwdx", "valid_x",
               "dhdy", "q_y", "Re_y", "b_y", "ub_y", "taub_y", "dpwdy", "valid_y"]
FIELD_ID = {n: i for i, n in enumerate(FIELD_NAMES)}
N_CENTER, N_XFACE = 21, 8

ABI_VERSION = 2  # SHK_ABI_VERSION of the header these structures follow

# Every exported entry point of include/shakti_hip.h; a load fails loudly if one is missing.
EXPORTS = ["shk_abi_version", "shk_last_error", "shk_create", "shk_destroy", "shk_set_solver_cfg", "shk_set_params",
           "shk_set_model", "shk_upload", "shk_download", "shk_set_initial_conditions", "shk_solve_elliptic",
           "shk_solve_parabolic", "shk_update_derived", "shk_picard_iteration", "shk_picard_loop", "shk_solve_log",
           "shk_parabolic_step", "shk_step_b", "shk_set_moulin_field", "shk_update_ieb", "shk_step", "shk_assemble",
           "shk_apply_operator", "shk_update_chebyshev_bounds", "shk_residual_norms", "shk_residual_norms_noalloc",
           "shk_apply_precond", "shk_download_system", "shk_set_sliding_coefficient", "shk_checkpoint_save",
           "shk_checkpoint_wait", "shk_checkpoint_load", "shk_snapshot", "shk_snapshot_fetch", "shk_field_stats", "shk_generate_glacier",
           "shk_time_kernel", "shk_profile_enable", "shk_profile_read", "shk_profile_reset"]
# ... and of include/shakti_learned.h
EXPORTS_LEARNED = ["shk_unet_create", "shk_unet_destroy", "shk_unet_last_error", "shk_unet_num_params",
                   "shk_unet_layer_info", "shk_unet_init_params", "shk_unet_set_params", "shk_unet_get_params",
                   "shk_unet_get_grads", "shk_unet_forward", "shk_unet_loss", "shk_unet_backward", "shk_unet_adam",
                   "shk_unet_train_step", "shk_unet_download_tensor", "shk_dev_alloc", "shk_dev_free",
                   "shk_dev_copy", "shk_attach_unet", "shk_set_learned_precond", "shk_record_pairs",
                   "shk_record_cg_pairs", "shk_recorded_pairs"]


class ShkGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("lx", C.c_double), ("ly", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dx2", C.c_double), ("dy2", C.c_double)]


class ShkParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ["rho_w", "rho_sw", "rho_i", "g", "nu", "n", "omega", "L", "br", "lr", "ct",
                                          "cw", "p_atm", "b_min", "b_max", "e_v"]]


class ShkModel(C.Structure):
    _fields_ = [("kfs", C.c_int32), ("gap_scheme", C.c_int32), ("sliding_law", C.c_int32),
                ("sliding_C", C.c_double), ("sliding_v0", C.c_double)]


class ShkGlacierSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64)] + [(n, C.c_double) for n in ["bed_amp", "corr_len", "trough_width",
                                                                    "trough_depth", "b0", "A_visc", "ub", "G",
                                                                    "ieb"]] + \
        [("n_moulins", C.c_int32), ("reserved", C.c_int32), ("q_min", C.c_double), ("q_max", C.c_double)]


class ShkSolverCfg(C.Structure):
    _fields_ = [("precond", C.c_int32), ("cheb_degree", C.c_int32), ("cheb_nsteps", C.c_int32), ("form", C.c_int32),
                ("itmax", C.c_int32), ("check_every", C.c_int32), ("atol", C.c_double), ("rtol", C.c_double),
                ("inner_rtol", C.c_double), ("cheb_bounds", C.c_int32), ("flags", C.c_int32)]


CHEB_BOUNDS = {"gershgorin": 0, "reference": 1}  # SHK_CHEB_BOUNDS_*


class ShkSolveStats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("solved", C.c_int32), ("zero_curvature", C.c_int32),
                ("bounds_warning", C.c_int32), ("refinements", C.c_int32), ("fallback", C.c_int32),
                ("rnorm0", C.c_double), ("rnorm", C.c_double), ("lambda_min", C.c_double),
                ("lambda_max", C.c_double), ("tired", C.c_int32), ("breakdown", C.c_int32),
                ("breakdown_value", C.c_double), ("lz_lambda_min", C.c_double), ("lz_lambda_max", C.c_double),
                ("gershgorin", C.c_double), ("true_residual", C.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class ShkUnetCfg(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_h", C.c_int32), ("max_w", C.c_int32), ("leaky_slope", C.c_float),
                ("loss_kind", C.c_int32)]


# enum shk_unet_tensor of include/shakti_learned.h: (name, channels, level)
UNET_TENSORS = [("x", 8, 0), ("a0", 16, 0), ("s0", 16, 0), ("p0", 16, 1), ("a1", 32, 1), ("s1", 32, 1),
                ("p1", 32, 2), ("a2", 64, 2), ("s2", 64, 2), ("p2", 64, 3), ("a3", 128, 3), ("t3", 128, 3),
                ("u2", 64, 2), ("v2", 64, 2), ("u1", 32, 1), ("v1", 32, 1), ("u0", 16, 0), ("v0", 16, 0),
                ("out", 1, 0), ("dout", 1, 0), ("dv0", 16, 0), ("du0", 16, 0), ("dcat0", 48, 0), ("dv1", 32, 1),
                ("du1", 32, 1), ("dcat1", 96, 1), ("dv2", 64, 2), ("du2", 64, 2), ("dcat2", 192, 2),
                ("dt3", 128, 3), ("da3", 128, 3), ("dp2", 64, 3), ("ds2", 64, 2), ("da2", 64, 2), ("dp1", 32, 2),
                ("ds1", 32, 1), ("da1", 32, 1), ("dp0", 16, 1), ("ds0", 16, 0), ("da0", 16, 0)]
UNET_TENSOR_ID = {n: i for i, (n, _, _) in enumerate(UNET_TENSORS)}
# enum shk_unet_layer
UNET_LAYERS = ["enc0a", "enc0b", "enc1a", "enc1b", "enc2a", "enc2b", "enc3a", "enc3b", "dec2a", "dec2b", "dec1a",
               "dec1b", "dec0a", "dec0b", "head"]


class ShaktiHIPError(RuntimeError):
    """A nonzero status from the C ABI (the reference raises ErrorException)."""


_lib = None


def load() -> C.CDLL:
    """Load libshakti_hip.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ShaktiHIPError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(str(LIB_PATH))
    for name in EXPORTS + EXPORTS_LEARNED:
        if not hasattr(L, name):
            raise ShaktiHIPError(f"{LIB_PATH} does not export {name}")
    vp, dp, i32, i64, sz, dbl = C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.c_int64, C.c_size_t, C.c_double
    L.shk_abi_version.restype = i32
    if L.shk_abi_version() != ABI_VERSION:
        raise ShaktiHIPError(f"{LIB_PATH} has ABI version {L.shk_abi_version()}, this binding expects {ABI_VERSION}; "
                             "rebuild it")
    L.shk_last_error.argtypes = [vp]
    L.shk_last_error.restype = C.c_char_p
    L.shk_create.argtypes = [C.POINTER(ShkGrid), C.POINTER(ShkParams), C.POINTER(ShkModel),
                             C.POINTER(ShkSolverCfg), C.POINTER(vp)]
    L.shk_destroy.argtypes = [vp]
    L.shk_set_solver_cfg.argtypes = [vp, C.POINTER(ShkSolverCfg)]
    L.shk_set_params.argtypes = [vp, C.POINTER(ShkParams)]
    L.shk_set_model.argtypes = [vp, C.POINTER(ShkModel)]
    L.shk_upload.argtypes = [vp, i32, dp, sz]
    L.shk_download.argtypes = [vp, i32, dp, sz]
    L.shk_set_initial_conditions.argtypes = [vp]
    L.shk_solve_elliptic.argtypes = [vp, C.POINTER(ShkSolveStats)]
    L.shk_solve_parabolic.argtypes = [vp, dbl, C.POINTER(ShkSolveStats)]
    L.shk_update_derived.argtypes = [vp]
    L.shk_picard_iteration.argtypes = [vp, dbl, C.POINTER(ShkSolveStats)]
    L.shk_picard_loop.argtypes = [vp, i32, dbl, i32, dbl, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)]
    L.shk_solve_log.argtypes = [vp, C.POINTER(ShkSolveStats), i32, C.POINTER(i32)]
    L.shk_parabolic_step.argtypes = [vp, dbl, C.POINTER(i64)]
    L.shk_step_b.argtypes = [vp, dbl]
    L.shk_set_moulin_field.argtypes = [vp, dp, sz]
    L.shk_update_ieb.argtypes = [vp, i32, dbl]
    L.shk_step.argtypes = [vp, dbl, i32, dbl, i32, dbl, i32, dbl, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)]
    L.shk_assemble.argtypes = [vp, i32, dbl]
    L.shk_apply_operator.argtypes = [vp, dp, dp]
    L.shk_update_chebyshev_bounds.argtypes = [vp, dp, dp, dp, C.POINTER(i32)]
    L.shk_residual_norms.argtypes = [vp, dp, dp]
    L.shk_residual_norms_noalloc.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp]
    L.shk_set_sliding_coefficient.argtypes = [vp, dp, sz]
    L.shk_checkpoint_save.argtypes = [vp, C.c_char_p, i64, dbl, i32]
    L.shk_checkpoint_wait.argtypes = [vp]
    L.shk_checkpoint_load.argtypes = [vp, C.c_char_p, C.POINTER(i64), dp]
    L.shk_snapshot.argtypes = [vp, C.POINTER(i32), i32]
    L.shk_snapshot_fetch.argtypes = [vp, i32, dp, sz]
    L.shk_field_stats.argtypes = [vp, i32, dp]
    L.shk_generate_glacier.argtypes = [vp, C.POINTER(ShkGlacierSpec)]
    L.shk_apply_precond.argtypes = [vp, dp, dp]
    L.shk_download_system.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.shk_time_kernel.argtypes = [vp, i32, i32, dp, dp]
    L.shk_profile_enable.argtypes = [vp, i32]
    L.shk_profile_read.argtypes = [vp, i32, C.POINTER(C.c_char_p), C.POINTER(i64), dp, dp, C.POINTER(i32)]
    L.shk_profile_reset.argtypes = [vp]
    fp = C.POINTER(C.c_float)
    L.shk_unet_create.argtypes = [C.POINTER(ShkUnetCfg), C.POINTER(vp)]
    L.shk_unet_destroy.argtypes = [vp]
    L.shk_unet_last_error.argtypes = [vp]
    L.shk_unet_last_error.restype = C.c_char_p
    L.shk_unet_num_params.restype = i64
    L.shk_unet_layer_info.argtypes = [i32, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32)]
    L.shk_unet_init_params.argtypes = [vp, C.c_uint64]
    L.shk_unet_set_params.argtypes = [vp, fp, i64]
    L.shk_unet_get_params.argtypes = [vp, fp, i64]
    L.shk_unet_get_grads.argtypes = [vp, fp, i64]
    L.shk_unet_forward.argtypes = [vp, vp, i32, i32, i32, vp]
    L.shk_unet_loss.argtypes = [vp, vp, i32, dp]
    L.shk_unet_backward.argtypes = [vp, vp]
    L.shk_unet_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.shk_unet_train_step.argtypes = [vp, vp, vp, i32, i32, i32, C.c_float, dp]
    L.shk_unet_download_tensor.argtypes = [vp, i32, fp, i64]
    L.shk_dev_alloc.argtypes = [sz, C.POINTER(vp)]
    L.shk_dev_free.argtypes = [vp]
    L.shk_dev_copy.argtypes = [vp, vp, sz, i32]
    L.shk_attach_unet.argtypes = [vp, vp]
    L.shk_set_learned_precond.argtypes = [vp, i32, dbl]
    L.shk_record_pairs.argtypes = [vp, vp, vp, i32]
    L.shk_record_cg_pairs.argtypes = [vp, vp, vp, i32]
    L.shk_recorded_pairs.argtypes = [vp, C.POINTER(i32)]
    _lib = L
    return L


def check_net(rc: int, net=None, what: str = "") -> None:
    """Like check() for a network handle (shk_unet_last_error)."""
    if rc != SHK_OK:
        msg = ""
        if net is not None:
            raw = load().shk_unet_last_error(net)
            msg = raw.decode() if raw else ""
        raise ShaktiHIPError(f"{what}: {ERRORS.get(rc, rc)} {msg}".strip())


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != SHK_OK:
        msg = ""
        if ctx is not None:
            raw = load().shk_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise ShaktiHIPError(f"{what}: {ERRORS.get(rc, rc)} {msg}".strip())
