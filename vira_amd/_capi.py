"""ctypes binding of the C-ABI in include/vira_gpu.h.

The shared library is the only compute path: if it is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

NUM_BINS = 8
MAX_LIGHTS = 8

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvira_gpu.so"

STATUS = {
    0: "VIRA_GPU_OK",
    1: "VIRA_GPU_ERR_INVALID",
    2: "VIRA_GPU_ERR_HIP",
    3: "VIRA_GPU_ERR_OOM",
    4: "VIRA_GPU_ERR_STATE",
    5: "VIRA_GPU_ERR_UNSUPPORTED",
}


class ViraGPUError(RuntimeError):
    """A non-OK status from the C-ABI (the reference throws std::runtime_error)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("stream", C.c_void_p)]


class Projection(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("spacing", C.c_double),
        ("radius", C.c_double),
        ("center", C.c_double * 3),
        ("offset", C.c_double * 3),
    ]


class BVHOpts(C.Structure):
    _fields_ = [("max_leaf", C.c_int), ("bins", C.c_int), ("width", C.c_int), ("traversal_cost", C.c_float),
                ("method", C.c_int), ("compress", C.c_int)]


class BVHStats(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_uint64),
        ("num_wide_nodes", C.c_uint64),
        ("width", C.c_uint32),
        ("wide_depth", C.c_uint32),
        ("num_leaves", C.c_uint64),
        ("max_depth", C.c_uint32),
        ("max_leaf_size", C.c_uint32),
        ("sah_cost", C.c_double),
        ("build_ms", C.c_double),
        ("bytes", C.c_uint64),
        ("method", C.c_int32),
        ("wide_sah_cost", C.c_double),
        ("compressed_bytes", C.c_uint64),
    ]


class Light(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("position", C.c_double * 3),
        ("radius", C.c_float),
        ("value", C.c_float * NUM_BINS),
    ]


class Material(C.Structure):
    _fields_ = [
        ("albedo", C.c_float * NUM_BINS),
        ("texture_dev", C.c_void_p),
        ("tex_width", C.c_int),
        ("tex_height", C.c_int),
        ("smooth_shading", C.c_int),
        ("tex_channels", C.c_int),
        ("normal_map_dev", C.c_void_p),
        ("nmap_width", C.c_int),
        ("nmap_height", C.c_int),
        ("material_id", C.c_uint32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("width", C.c_int),
        ("height", C.c_int),
        ("kinv", C.c_float * 6),
        ("rot", C.c_float * 9),
        ("position", C.c_double * 3),
        ("z_dir", C.c_float),
        ("view_normal", C.c_float * 9),
    ]


class RenderOpts(C.Structure):
    _fields_ = [
        ("samples", C.c_int),
        ("bounces", C.c_int),
        ("russian_roulette", C.c_int),
        ("rr_start", C.c_int),
        ("seed", C.c_uint32),
        ("integrator", C.c_int),
        ("show_background", C.c_int),
        ("adaptive_sampling", C.c_int),
        ("samples_per_batch", C.c_int),
        ("sampling_tolerance", C.c_float),
        ("samples_to_detect_miss", C.c_int),
        ("count_visits", C.c_int),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("radiance_dev", C.c_void_p),
        ("direct_dev", C.c_void_p),
        ("indirect_dev", C.c_void_p),
        ("albedo_dev", C.c_void_p),
        ("normal_dev", C.c_void_p),
        ("depth_dev", C.c_void_p),
        ("alpha_dev", C.c_void_p),
        ("tri_id_dev", C.c_void_p),
        ("instance_id_dev", C.c_void_p),
        ("normal_camera_dev", C.c_void_p),
        ("mesh_id_dev", C.c_void_p),
        ("material_id_dev", C.c_void_p),
    ]


class VisitCounts(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint64),
        ("leaves", C.c_uint64),
        ("triangles", C.c_uint64),
        ("stack_spills", C.c_uint64),
        ("tlas_nodes", C.c_uint64),
        ("max_stack", C.c_uint32),
        ("wave_node_steps", C.c_uint64),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("bounce_rays", C.c_uint64),
        ("total_ms", C.c_double),
        ("trace_closest_ms", C.c_double),
        ("trace_shadow_ms", C.c_double),
        ("shade_ms", C.c_double),
        ("trace_closest_launches", C.c_uint32),
        ("trace_shadow_launches", C.c_uint32),
        ("integrator", C.c_int),
        ("fused_ms", C.c_double),
        ("fused_launches", C.c_uint32),
        ("closest_hits", C.c_uint64),
        ("visits_counted", C.c_int),
        ("closest_visits", VisitCounts),
        ("shadow_visits", VisitCounts),
    ]


def struct_dict(st) -> dict:
    """Fields of a ctypes structure as a dict (nested structures as dicts)."""
    out = {}
    for f, _ in st._fields_:
        v = getattr(st, f)
        out[f] = struct_dict(v) if isinstance(v, C.Structure) else v
    return out


class DIOpts(C.Structure):
    _fields_ = [("samples", C.c_int), ("seed", C.c_uint32), ("record_budget", C.c_uint64)]


class DIStats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("ms", C.c_double),
        ("trace_ms", C.c_double),
        ("used_records", C.c_int32),
        ("record_bytes", C.c_uint64),
    ]


class Request(C.Structure):
    _fields_ = [
        ("mesh_id", C.c_int32),
        ("camera", Camera),
        ("opts", RenderOpts),
        ("num_lights", C.c_int32),
        ("lights", Light * MAX_LIGHTS),
        ("out", Outputs),
    ]


class BatchConfig(C.Structure):
    _fields_ = [("bvh_width", C.c_int32), ("waves_per_block", C.c_int32), ("waves_per_cu", C.c_int32),
                ("tiles_per_grab", C.c_int32), ("order", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("bounce_rays", C.c_uint64),
        ("tiles", C.c_uint64),
        ("ms", C.c_double),
        ("graph_reused", C.c_int32),
        ("closest_hits", C.c_uint64),
    ]


class Sensor(C.Structure):
    _fields_ = [("optical_efficiency", C.c_float * NUM_BINS), ("aperture_area", C.c_float)]


class PSF(C.Structure):
    _fields_ = [("kind", C.c_int), ("sigma_x", C.c_float * NUM_BINS), ("sigma_y", C.c_float * NUM_BINS),
                ("angle_deg", C.c_float), ("focal_length", C.c_float), ("aperture_diameter", C.c_float),
                ("pixel_size", C.c_float * 2), ("wavelength", C.c_float * NUM_BINS), ("supersample", C.c_int)]


class DenoiseOpts(C.Structure):
    _fields_ = [("max_levels_direct", C.c_int32), ("max_levels_indirect", C.c_int32),
                ("depth_threshold", C.c_float), ("normal_threshold", C.c_float),
                ("min_normal_weight", C.c_float), ("epsilon", C.c_float)]


class SensorSim(C.Structure):
    _fields_ = [
        ("exposure_time", C.c_float),
        ("photon_energy", C.c_float * NUM_BINS),
        ("quantum_efficiency", C.c_float * NUM_BINS),
        ("gain", C.c_float),
        ("max_adu", C.c_float),
        ("photon_noise", C.c_int32),
        ("noise_model", C.c_int32),
        ("dark_current", C.c_float),
        ("readout_noise_mean", C.c_float),
        ("readout_noise_std", C.c_float),
        ("horizontal_scale", C.c_float),
        ("vertical_scale", C.c_float),
        ("horizontal_pattern_period", C.c_int32),
        ("vertical_pattern_period", C.c_int32),
        ("seed", C.c_uint32),
        ("filter_mosaic_dev", C.c_void_p),
    ]


class StarPSF(C.Structure):
    _fields_ = [("count", C.c_int32), ("size", C.c_int32 * 4), ("kernel_dev", C.c_void_p * 4),
                ("edge_max", (C.c_float * NUM_BINS) * 4), ("minimum_power", C.c_float),
                ("intrinsic", C.c_float * 6), ("cull_behind", C.c_int32)]


class Star(C.Structure):
    """One star quipu record: the reference's Star object as it lies in memory
    (star.hpp:27-38): initialized flag, padding, then RA, DE, the two proper
    motions (rad/yr), V magnitude, temperature (K) and solid angle (sr)."""
    _fields_ = [("initialized", C.c_uint8), ("pad", C.c_uint8 * 7), ("ra", C.c_double), ("de", C.c_double),
                ("pm_ra", C.c_double), ("pm_de", C.c_double), ("vmag", C.c_double),
                ("temperature", C.c_double), ("omega", C.c_double)]


class QldInfo(C.Structure):
    _fields_ = [("levels", C.c_int32), ("has_albedos", C.c_int32), ("transformation", C.c_double * 16),
                ("normal", C.c_float * 3), ("cone_angle", C.c_float)]


class QldReadOpts(C.Structure):
    _fields_ = [("read_albedos", C.c_int32), ("default_albedo", C.c_float), ("albedo_profile", C.c_float * NUM_BINS),
                ("offset", C.c_double * 3)]


class QldLevel(C.Structure):
    _fields_ = [("level", C.c_int32), ("gsd", C.c_double), ("width", C.c_int32), ("height", C.c_int32),
                ("height_bits", C.c_int32), ("compressed", C.c_int32), ("albedo_type", C.c_int32),
                ("is_ogr", C.c_int32), ("read_ms", C.c_double), ("decode_ms", C.c_double), ("build_ms", C.c_double)]


class Instance(C.Structure):
    _fields_ = [("mesh_id", C.c_int32), ("transform", C.c_double * 16)]


# (name, restype, argtypes) for every entry point of include/vira_gpu.h
_P = C.c_void_p
_S = C.c_int
SIGNATURES = [
    ("vira_gpu_create", _S, [C.POINTER(_P), C.POINTER(Opts)]),
    ("vira_gpu_destroy", _S, [_P]),
    ("vira_gpu_set_stream", _S, [_P, _P]),
    ("vira_gpu_synchronize", _S, [_P]),
    ("vira_gpu_last_error", C.c_char_p, []),
    ("vira_gpu_version", C.c_char_p, []),
    ("vira_gpu_synth_fbm", _S, [_P, C.c_int, C.c_int, C.c_uint32, C.c_double, C.c_int, C.c_double, _P]),
    ("vira_gpu_synth_vertex_albedo", _S, [_P, C.c_size_t, C.c_uint32, _P]),
    ("vira_gpu_dem_mesh", _S, [_P, _P, C.c_int, C.c_int, C.POINTER(Projection), _P, C.POINTER(C.c_int32)]),
    ("vira_gpu_mesh_upload", _S, [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P, C.POINTER(C.c_int32)]),
    ("vira_gpu_mesh_info", _S, [_P, C.c_int32, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("vira_gpu_mesh_download", _S, [_P, C.c_int32, _P, _P, _P]),
    ("vira_gpu_mesh_release", _S, [_P, C.c_int32]),
    ("vira_gpu_blas_build", _S, [_P, C.c_int32, C.POINTER(BVHOpts)]),
    ("vira_gpu_blas_stats", _S, [_P, C.c_int32, C.POINTER(BVHStats)]),
    ("vira_gpu_blas_refit", _S, [_P, C.c_int32]),
    ("vira_gpu_tlas_build", _S, [_P, C.POINTER(Instance), C.c_size_t]),
    ("vira_gpu_lights_set", _S, [_P, C.POINTER(Light), C.c_int]),
    ("vira_gpu_material_set", _S, [_P, C.c_int32, C.POINTER(Material)]),
    ("vira_gpu_environment_set", _S, [_P, C.POINTER(C.c_float), _P, C.c_int, C.c_int]),
    ("vira_gpu_render", _S, [_P, C.POINTER(Camera), C.POINTER(RenderOpts), C.POINTER(Outputs), C.POINTER(RenderStats)]),
    ("vira_gpu_render_noalloc", _S, [_P, C.POINTER(Camera), C.POINTER(RenderOpts), C.POINTER(Outputs),
                                     C.POINTER(RenderStats)]),
    ("vira_gpu_trace_rays", _S, [_P, C.c_int32, _P, C.c_size_t, C.c_int, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    ("vira_gpu_dem_update", _S, [_P, C.c_int32, _P]),
    ("vira_gpu_dem_update_refit", _S, [_P, C.c_int32, _P]),
    ("vira_gpu_di_render", _S, [_P, C.POINTER(Camera), C.POINTER(DIOpts), _P, C.POINTER(DIStats)]),
    ("vira_gpu_adjoint", _S, [_P, C.POINTER(Camera), C.POINTER(DIOpts), _P, _P, _P, C.POINTER(DIStats)]),
    ("vira_gpu_received_power", _S, [_P, C.POINTER(Camera), C.POINTER(Sensor), _P, _P, _P]),
    ("vira_gpu_psf_kernel", _S, [_P, C.POINTER(PSF), C.c_int, _P]),
    ("vira_gpu_convolve", _S, [_P, _P, C.c_int, C.c_int, _P, C.c_int, _P]),
    ("vira_gpu_simulate_sensor", _S, [_P, C.POINTER(SensorSim), _P, C.c_int, C.c_int, _P, _P]),
    ("vira_gpu_lz4_decode_block", C.c_int64, [_P, C.c_size_t, _P, C.c_size_t]),
    ("vira_gpu_qld_open", _S, [C.c_char_p, C.POINTER(_P), C.POINTER(QldInfo)]),
    ("vira_gpu_qld_close", _S, [_P]),
    ("vira_gpu_qld_level_gsd", _S, [_P, C.c_int32, C.POINTER(C.c_double)]),
    ("vira_gpu_qld_select", _S, [_P, C.c_double, C.POINTER(C.c_int32)]),
    ("vira_gpu_qld_read", _S, [_P, _P, C.c_int32, C.POINTER(QldReadOpts), C.POINTER(C.c_int32), _P, _P,
                               C.POINTER(QldLevel)]),
    ("vira_gpu_denoise", _S, [_P, C.POINTER(DenoiseOpts), C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
    ("vira_gpu_render_stars", _S, [_P, C.POINTER(Camera), C.POINTER(Sensor), C.POINTER(StarPSF), _P, _P,
                                   C.c_size_t, _P]),
    ("vira_gpu_star_light", _S, [_P, _P, C.c_size_t, C.c_double, _P, _P]),
    ("vira_gpu_qsc_read", _S, [C.c_char_p, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("vira_gpu_batch_submit", _S, [_P, C.POINTER(Request), C.c_size_t, C.POINTER(BatchConfig),
                                   C.POINTER(BatchStats)]),
    ("vira_gpu_batch_autotune", _S, [_P, C.POINTER(Request), C.c_size_t, C.c_int, C.POINTER(BatchConfig),
                                     C.POINTER(C.c_double)]),
]

_lib = None


def lib_path() -> Path:
    return _LIB_PATH


def load() -> C.CDLL:
    """Loads libvira_gpu.so (built by vira_amd/build.py); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(
            f"{_LIB_PATH} is missing: build it with `python -m vira_amd.build` "
            "(the HIP library is the only compute path)"
        )
    lib = C.CDLL(str(_LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().vira_gpu_last_error().decode(errors="replace")
        raise ViraGPUError(status, msg)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return t.data_ptr()
