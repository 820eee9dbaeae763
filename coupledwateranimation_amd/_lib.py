"""ctypes binding of libcwa.so (the C ABI in include/cwa.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CWA_LIB", PKG / "_build" / "libcwa.so"))

CWA_OK = 0
CWA_EINVAL = -1
CWA_EHIP = -2
CWA_ENOMEM = -3
CWA_ESTATE = -4
CWA_ENONFINITE = -5
CWA_NUM_KERNELS = 10

f3 = C.c_float * 3
f2 = C.c_float * 2
i3 = C.c_int32 * 3


class Params(C.Structure):
    """cwa_params (include/cwa.h)."""

    _fields_ = [
        ("mass", C.c_float), ("h", C.c_float), ("visc", C.c_float), ("rest_rho", C.c_float),
        ("gas_const", C.c_float), ("particle_radius", C.c_float), ("spiky_coeff", C.c_float),
        ("visc_coeff", C.c_float), ("gravity", f3), ("crest_threshold", C.c_float),
        ("breaking_mass", C.c_float), ("breaking_visc", C.c_float), ("torque_coeff", C.c_float),
        ("drag_coeff", C.c_float), ("wave_force_coeff", C.c_float), ("wave_vel_du", C.c_float),
        ("vel_damping", C.c_float), ("bounce", C.c_float), ("foam_speed", C.c_float),
        ("foam_force", C.c_float), ("foam_rho", C.c_float), ("foam_pressure", C.c_float),
        ("foam_vel", C.c_float), ("upper", f3), ("lower", f3), ("ceiling_extra", C.c_float),
        ("wave_origin", f2), ("wave_uv_scale", f2), ("lambda_x", C.c_float),
        ("lambda_z", C.c_float), ("atten", C.c_float), ("beta", C.c_float),
        ("wave_type", C.c_float), ("wake_threshold", C.c_float), ("wake_amount", C.c_float),
        ("couple_band", C.c_float), ("couple_strength", C.c_float), ("grid_min", f3),
        ("cell_size", C.c_float), ("grid_dims", i3), ("grid_ydiv", C.c_int32),
        ("check_finite", C.c_int32),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, C.Array) else v
        return out


class Init(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("pos", C.POINTER(C.c_float)), ("vel", C.POINTER(C.c_float)),
        ("force", C.POINTER(C.c_float)), ("wave_w", C.c_int32), ("wave_h", C.c_int32),
        ("height0", C.POINTER(C.c_float)), ("height1", C.POINTER(C.c_float)),
    ]


class ParticleView(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("pos", C.c_void_p), ("vel", C.c_void_p), ("force", C.c_void_p),
        ("rho", C.c_void_p), ("prs", C.c_void_p), ("id", C.c_void_p),
    ]


class GridView(C.Structure):
    _fields_ = [
        ("dims", i3), ("num_cells", C.c_int32), ("n", C.c_int32), ("key", C.c_void_p),
        ("count", C.c_void_p), ("offset", C.c_void_p), ("list", C.c_void_p),
        ("flux", C.c_void_p),
    ]


class HostState(C.Structure):
    _fields_ = [
        ("pos", C.POINTER(C.c_float)), ("vel", C.POINTER(C.c_float)),
        ("force", C.POINTER(C.c_float)), ("rho", C.POINTER(C.c_float)),
        ("prs", C.POINTER(C.c_float)), ("height0", C.POINTER(C.c_float)),
        ("height1", C.POINTER(C.c_float)),
    ]


class KernelTimes(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("ms", C.c_float * CWA_NUM_KERNELS),
        ("launches", C.c_int32 * CWA_NUM_KERNELS),
    ]


class Grad(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("visc", "gas_const", "rest_rho", "lambda_x",
                                          "lambda_z", "atten", "couple_strength", "beta")]


STRUCTS = [Params, Init, ParticleView, GridView, HostState, KernelTimes, Grad]

# --- batched 2D step (include/cwa2d.h) -------------------------------------
CWA2D_NUM_KERNELS = 7


class Params2D(C.Structure):
    """cwa2d_params (include/cwa2d.h)."""

    _fields_ = [
        ("h", C.c_float), ("rest_rho", C.c_float), ("gas_const", C.c_float),
        ("visc", C.c_float), ("psi", C.c_float), ("gravity", f2), ("lo", f2), ("hi", f2),
        ("walls", C.c_float * 4), ("enable_vel", C.c_float), ("wave_ref", C.c_float),
        ("wave_press", C.c_float), ("wave_dy", C.c_float), ("wave_maskv", C.c_float),
        ("atten", C.c_float), ("beta", C.c_float), ("couple_band", C.c_float),
        ("couple_strength", C.c_float), ("sph_substeps", C.c_int32),
        ("check_finite", C.c_int32),
    ]

    def copy(self) -> "Params2D":
        p = Params2D()
        C.pointer(p)[0] = self
        return p


CWA2D_WAVE_LEAPFROG = 0
CWA2D_WAVE_SHALLOW = 1


class Scene2D(C.Structure):
    _fields_ = [("n", C.c_int32), ("wave_len", C.c_int32), ("wave_substeps", C.c_int32),
                ("wave_model", C.c_int32), ("mass", C.c_float), ("lambda_", C.c_float),
                ("wave_dx", C.c_float)]


class Tuning2D(C.Structure):
    _fields_ = [("block_density", C.c_int32), ("block_force", C.c_int32),
                ("grid_density", C.c_int32), ("grid_force", C.c_int32),
                ("xcd_queues", C.c_int32), ("wave_block", C.c_int32),
                ("ms_per_step", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Views2D(C.Structure):
    _fields_ = [
        ("nscenes", C.c_int32), ("n_total", C.c_int32), ("wave_total", C.c_int32),
        ("grid_nx", C.c_int32), ("grid_ny", C.c_int32), ("cells_total", C.c_int32),
        ("xv", C.c_void_p), ("rp", C.c_void_p), ("idf", C.c_void_p), ("key", C.c_void_p),
        ("offset", C.c_void_p), ("h0", C.c_void_p), ("h1", C.c_void_p),
        ("scene_begin", C.c_void_p), ("wave_begin", C.c_void_p),
    ]


class KernelTimes2D(C.Structure):
    _fields_ = [("ms", C.c_float * CWA2D_NUM_KERNELS)]


STRUCTS_2D = [Params2D, Scene2D, Tuning2D, Views2D, KernelTimes2D]

# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "cwa_abi_version": (C.c_int, []),
    "cwa_struct_size": (C.c_int, [C.c_int]),
    "cwa_device_count": (C.c_int, []),
    "cwa_last_error": (C.c_char_p, []),
    "cwa_kernel_name": (C.c_char_p, [C.c_int]),
    "cwa_scene_particles_written": (C.c_int, [C.c_void_p]),
    "cwa_scene_prepare_graph": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(C.c_int)]),
    "cwa_set_lane_split_limit": (C.c_int, [C.c_longlong, C.POINTER(C.c_longlong)]),
    "cwa_params_reference": (None, [C.POINTER(Params)]),
    "cwa_scene_create": (C.c_int, [C.POINTER(Params), C.POINTER(Init), C.POINTER(C.c_void_p)]),
    "cwa_scene_reinit": (C.c_int, [C.c_void_p, C.POINTER(Init)]),
    "cwa_scene_destroy": (C.c_int, [C.c_void_p]),
    "cwa_scene_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "cwa_scene_get_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "cwa_scene_wave_reinit": (C.c_int, [C.c_void_p]),
    "cwa_scene_wave_reinit_from_texture": (
        C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.c_int]),
    "cwa_scene_wave_compute": (C.c_int, [C.c_void_p, C.c_float]),
    "cwa_scene_height": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cwa_scene_step": (C.c_int, [C.c_void_p, C.c_float, C.c_int]),
    "cwa_scene_step_profiled": (C.c_int, [C.c_void_p, C.c_float, C.c_int,
                                          C.POINTER(KernelTimes)]),
    "cwa_scene_step_graph": (C.c_int, [C.c_void_p, C.c_float, C.c_int]),
    "cwa_scene_sync": (C.c_int, [C.c_void_p]),
    "cwa_scene_stream": (C.c_void_p, [C.c_void_p]),
    "cwa_scene_particles": (C.c_int, [C.c_void_p, C.POINTER(ParticleView)]),
    "cwa_scene_grid": (C.c_int, [C.c_void_p, C.POINTER(GridView)]),
    "cwa_scene_readback": (C.c_int, [C.c_void_p, C.POINTER(HostState)]),
    "cwa_scene_neighbour_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "cwa_scan_exclusive_i32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "cwa_grid_build": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_float,
                                 C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    "cwa_grid_build_static": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_float,
                                        C.POINTER(C.c_int32), C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
    "cwa_grid_build_boxes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int64,
                                       C.POINTER(C.c_int64), C.c_void_p]),
    "cwa_tape_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "cwa_tape_forward": (C.c_int, [C.c_void_p, C.c_float]),
    "cwa_tape_backward": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(Grad),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "cwa_tape_destroy": (C.c_int, [C.c_void_p]),
    "cwa_gen_lattice": (C.c_int, [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float,
                                  C.POINTER(C.c_int32), C.c_uint64]),
    "cwa_gen_bumps": (C.c_int, [C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_float, C.c_float, C.c_float, C.c_uint64]),
    "cwa_fit_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p),
                                 C.POINTER(C.POINTER(C.c_float)), C.c_int32, C.c_int32,
                                 C.c_float, C.POINTER(C.c_void_p)]),
    "cwa_fit_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(Grad)]),
    "cwa_fit_tape_bytes": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    "cwa_fit_destroy": (C.c_int, [C.c_void_p]),
    # batched 2D (include/cwa2d.h)
    "cwa2d_abi_version": (C.c_int, []),
    "cwa2d_struct_size": (C.c_int, [C.c_int]),
    "cwa2d_kernel_name": (C.c_char_p, [C.c_int]),
    "cwa2d_params_reference": (C.c_int, [C.POINTER(Params2D)]),
    "cwa2d_wave_init_reference": (C.c_int, [C.c_int32, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float)]),
    "cwa2d_wave_init_shallow": (C.c_int, [C.c_int32, C.c_float, C.POINTER(C.c_float),
                                          C.POINTER(C.c_float)]),
    "cwa2d_batch_splash": (C.c_int, [C.c_void_p, C.c_int32, C.c_float]),
    "cwa2d_batch_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cwa2d_batch_create": (C.c_int, [C.POINTER(Params2D), C.c_int32, C.POINTER(Scene2D),
                                     C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_void_p)]),
    "cwa2d_batch_destroy": (C.c_int, [C.c_void_p]),
    "cwa2d_batch_reinit": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "cwa2d_batch_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params2D)]),
    "cwa2d_batch_get_params": (C.c_int, [C.c_void_p, C.POINTER(Params2D)]),
    "cwa2d_batch_step": (C.c_int, [C.c_void_p, C.c_float, C.c_int32]),
    "cwa2d_batch_step_graph": (C.c_int, [C.c_void_p, C.c_float, C.c_int32]),
    "cwa2d_batch_step_profiled": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(KernelTimes2D)]),
    "cwa2d_batch_sync": (C.c_int, [C.c_void_p]),
    "cwa2d_batch_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "cwa2d_batch_views": (C.c_int, [C.c_void_p, C.POINTER(Views2D)]),
    "cwa2d_batch_readback": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    "cwa2d_batch_tune": (C.c_int, [C.c_void_p, C.c_float, C.c_int32, C.POINTER(Tuning2D)]),
    "cwa2d_batch_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning2D)]),
    "cwa2d_batch_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning2D)]),
}

_lib = None


class CwaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cwa error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH} is missing "
                "(run `python -m coupledwateranimation_amd.build`)")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Loading torch first makes libcwa bind to that
        # one, so the process has a single HIP runtime; loading libcwa first
        # would map a second runtime and torch would then see no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.cwa_abi_version() != 1:
            raise RuntimeError("libcwa ABI version mismatch")
        for k, st in enumerate(STRUCTS):
            if L.cwa_struct_size(k) != C.sizeof(st):
                raise RuntimeError(f"struct {st.__name__} size mismatch with libcwa")
        if L.cwa2d_abi_version() != 2:
            raise RuntimeError("libcwa 2D ABI version mismatch")
        for k, st in enumerate(STRUCTS_2D):
            if L.cwa2d_struct_size(k) != C.sizeof(st):
                raise RuntimeError(f"struct {st.__name__} size mismatch with libcwa")
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != CWA_OK:
        raise CwaError(rc, lib().cwa_last_error().decode())


def iptr(a):
    """int32* of a C-contiguous int32 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def fptr(a):
    """float* of a C-contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))
