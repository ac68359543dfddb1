"""ctypes binding of libnfx.so (declarations in include/nfx.h).

The product path has no fallback: if the library is missing, or no HIP device is present, the
calls raise.  Device buffers are torch tensors; only their ``data_ptr()`` crosses the boundary.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnfx.so")

NFX_SVKK, NFX_SVK, NFX_MR, NFX_VISCOPLASTIC = 0, 1, 2, 3
NFX_LINEAR_ELASTIC, NFX_LINEAR_DEVIATORIC, NFX_NEO_HOOKEAN_2, NFX_PORO_NEO_HOOKEAN = 4, 5, 6, 7
SCATTER_ATOMIC, SCATTER_CSR, SCATTER_TILED = 0, 1, 2
INDENTER_NONE, INDENTER_SPHERE, INDENTER_PLANE = 0, 1, 2

c_i32, c_i64, c_f64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class NfxError(RuntimeError):
    pass


class Elastic(ctypes.Structure):
    _fields_ = [("model", c_i32), ("E", c_f64), ("nu", c_f64), ("ratio", c_f64), ("shear", c_f64), ("bulk", c_f64),
                ("shear2", c_f64), ("volume_fraction", c_f64)]


class Viscoelastic(ctypes.Structure):
    _fields_ = [("E", c_f64), ("nu", c_f64), ("nterms", c_i32), ("weights", c_f64 * 8), ("taus", c_f64 * 8)]


class Viscoplastic(ctypes.Structure):
    _fields_ = [
        ("mu1", c_f64),
        ("mu2", c_f64),
        ("kappa", c_f64),
        ("flow_exponent", c_f64),
        ("yield_stress", c_f64),
        ("flow_rate", c_f64),
        ("hardening_threshold", c_f64),
        ("hardening_exponent", c_f64),
        ("yield_offset", c_f64),
    ]


class Plan(ctypes.Structure):
    _fields_ = [
        ("N", c_i64),
        ("T", c_i64),
        ("tetras", c_vp),
        ("dminv", c_vp),
        ("v0", c_vp),
        ("ref", c_vp),
        ("disp", c_vp),
        ("vel", c_vp),
        ("force", c_vp),
        ("fold", c_vp),
        ("model", c_i32),
        ("E", c_f64),
        ("nu", c_f64),
        ("ratio", c_f64),
        ("vp", Viscoplastic),
        ("vp_state", c_vp),
        ("epsbar", c_vp),
        ("von_mises", c_vp),
        ("pressure", c_vp),
        ("scatter", c_i32),
        ("csr_offsets", c_vp),
        ("csr_slots", c_vp),
        ("corners", c_vp),
        ("tile_k", c_i32),
        ("tile_max_nodes", c_i32),
        ("tile_count", c_vp),
        ("tile_nodes", c_vp),
        ("tile_lend", c_vp),
        ("tile_corners", c_vp),
        ("tile_dst", c_vp),
        ("node_off", c_vp),
        ("partials", c_vp),
        ("substrate", c_i32),
        ("substrate_k", c_f64),
        ("wall", c_f64),
        ("indenter", c_i32),
        ("ind_k", c_f64),
        ("ind_radius", c_f64),
        ("ind_start", c_f64),
        ("ind_speed", c_f64),
        ("ind_move_steps", c_f64),
        ("fix_top_bottom", c_i32),
        ("shear", c_f64),
        ("bulk", c_f64),
        ("shear2", c_f64),
        ("volume_fraction", c_f64),
        ("vp_sym", c_i32),
        ("stream_nt", c_i32),
        ("shape_from_ref", c_i32),
        ("vp_active", c_vp),
        ("disp_b", c_vp),
        ("vel_b", c_vp),
        ("partials_b", c_vp),
        ("tile_range", c_vp),
    ]


class Batch(ctypes.Structure):
    _fields_ = [
        ("ncells", c_i32),
        ("tile_ntet", c_vp),
        ("tile_cell", c_vp),
        ("node_cell", c_vp),
        ("cell_node_off", c_vp),
        ("cells", c_vp),
        ("curves", c_vp),
        ("ncurve", c_i32),
        ("curve_steps", c_i64),
        ("tile_nodes16", c_vp),
        ("tile_dst16", c_vp),
        ("tile_node_base", c_vp),
        ("tile_dst_base", c_vp),
    ]


TILE_MAX_NODES = 256  # tile-node index in 8 bits (nfx_plan.tile_corners)

NFX_MLP_STORE, NFX_MLP_FWD_GELU, NFX_MLP_BWD_GELU, NFX_MLP_ADAM = 0, 1, 2, 3


class MlpGemm(ctypes.Structure):
    _fields_ = [
        ("epilogue", c_i32),
        ("M", c_i32),
        ("N", c_i32),
        ("K", c_i32),
        ("A", c_vp),
        ("lda", c_i32),
        ("B", c_vp),
        ("ldb", c_i32),
        ("bias", c_vp),
        ("C", c_vp),
        ("Z", c_vp),
        ("H", c_vp),
        ("HT", c_vp),
        ("adam_m", c_vp),
        ("adam_v", c_vp),
        ("lr_t", ctypes.c_float),
        ("bc2", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("grad_scale", ctypes.c_float),
        ("workspace", c_vp),
        ("split_k", c_i32),
    ]


# name -> (restype, argtypes); every symbol include/nfx.h declares
SIGNATURES = {
    "nfx_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "nfx_init": (ctypes.c_int, [ctypes.c_int, c_vp, ctypes.POINTER(c_vp)]),
    "nfx_destroy": (ctypes.c_int, [c_vp]),
    "nfx_stream_sync": (ctypes.c_int, [c_vp]),
    "nfx_last_error": (ctypes.c_char_p, []),
    "nfx_device_name": (ctypes.c_int, [c_vp, ctypes.c_char_p, c_i32]),
    "nfx_precompute_f64": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "nfx_tetra_force": (ctypes.c_int, [c_vp, ctypes.POINTER(Elastic), c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "nfx_tetra_corner_forces": (ctypes.c_int,
                                [c_vp, ctypes.POINTER(Elastic), c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "nfx_viscoplastic_force": (ctypes.c_int, [c_vp, ctypes.POINTER(Viscoplastic), c_vp, c_vp, c_vp, c_i64, c_i64,
                                              c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "nfx_gather_corners": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "nfx_vp_q_from_fp": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "nfx_vp_fp_from_q": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "nfx_vp_tile_flags": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan)]),
    "nfx_points_from_disp": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "nfx_disp_from_points": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "nfx_graph_captures": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i64)]),
    "nfx_round_tip_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64, c_f64]),
    "nfx_linear_viscoelastic_force": (ctypes.c_int, [c_vp, ctypes.POINTER(Viscoelastic), c_vp, c_vp, c_vp, c_i64,
                                                      c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "nfx_adhesive_plane_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64, c_f64, c_f64, c_vp]),
    "nfx_point_zero_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64]),
    "nfx_center_of_mass": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "nfx_com_force": (ctypes.c_int, [c_vp, c_vp, c_i64, c_f64, c_vp]),
    "nfx_over_damped": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_i32]),
    "nfx_velocity_verlet2": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_f64]),
    "nfx_volume_integral_force": (ctypes.c_int, [c_vp, c_i32, c_vp, c_vp, c_i64, c_i64, c_f64, c_f64, c_f64, c_f64,
                                                  c_f64, c_f64, c_vp, c_vp]),
    "nfx_adhesive_plane_surface_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_f64, c_f64, c_f64]),
    "nfx_hard_adhesive_plane_surface_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_f64, c_f64,
                                                              c_f64]),
    "nfx_legacy_force": (ctypes.c_int, [c_vp, ctypes.POINTER(Elastic), c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp,
                                         c_vp, c_vp]),
    "nfx_batch_geometry": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "nfx_build_node_csr_host": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_vp]),
    "nfx_build_tiles_host": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i32, ctypes.POINTER(c_i32), ctypes.POINTER(c_i64),
                                            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "nfx_sphere_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64, c_f64, c_f64]),
    "nfx_substrate_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64, c_f64]),
    "nfx_plane_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_f64, c_f64]),
    "nfx_velocity_verlet": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i32]),
    "nfx_tip_force": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_i32, c_f64, c_f64, c_f64, c_vp]),
    "nfx_tip_force_noalloc": (ctypes.c_int, [c_vp, c_vp, c_i64, c_i32, c_f64, c_f64, c_f64, c_vp, c_vp]),
    "nfx_run_steps": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan), c_i64, c_i64, c_vp, c_i32]),
    "nfx_node_update": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan), c_i64, c_vp]),
    "nfx_time_passes": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan), c_i64, c_i32, ctypes.POINTER(c_f64),
                                       ctypes.POINTER(c_f64)]),
    "nfx_mlp_gemm": (ctypes.c_int, [c_vp, ctypes.POINTER(MlpGemm)]),
    "nfx_mlp_mse": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp]),
    "nfx_mlp_bias_adam": (ctypes.c_int, [c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "nfx_stream_copy": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_i32, ctypes.POINTER(c_f64)]),
    "nfx_batch_run_steps": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan), ctypes.POINTER(Batch), c_i64, c_i64, c_i32]),
    "nfx_batch_time_passes": (ctypes.c_int, [c_vp, ctypes.POINTER(Plan), ctypes.POINTER(Batch), c_i64, c_i32,
                                             ctypes.POINTER(c_f64), ctypes.POINTER(c_f64), ctypes.POINTER(c_f64)]),
}

_lib = None
_extra_signatures = []


def _bind(L, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load libnfx.so (raises if it has not been built).

    torch is imported first: its HIP runtime (SONAME libamdhip64.so.7) is then already in the
    process when libnfx.so asks for libamdhip64.so.7, so both share one runtime.  Loaded the other
    way round, torch would bring a second HIP runtime of its own that finds no device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NfxError(f"{LIB_PATH} is missing: build it with `python -m nofluidx3d_amd.build`")
        import torch  # noqa: F401  (loads torch's HIP runtime; initialises no device)

        L = ctypes.CDLL(LIB_PATH)
        _bind(L, SIGNATURES)
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().nfx_last_error().decode(errors="replace")
        raise NfxError(f"{what}: {msg} (code {rc})" if what else f"{msg} (code {rc})")


def call(name: str, *args):
    """Call ``name`` and raise NfxError on a non-zero return.  The argument count is checked
    against the declared signature (ctypes would pass surplus arguments on silently)."""
    want = len(SIGNATURES[name][1])
    if len(args) != want:
        raise TypeError(f"{name} takes {want} arguments ({len(args)} given)")
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int:
    """Device address of a torch tensor (or 0 for None)."""
    if t is None:
        return None
    return t.data_ptr()
