"""ctypes binding of liblmx.so (declarations mirror include/lmx.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# LMX_LIBRARY selects another build of the library (for comparing builds).
LIB_PATH = os.environ.get("LMX_LIBRARY") or os.path.join(os.path.dirname(HERE), "liblmx.so")


class LmxError(RuntimeError):
    """A call into liblmx returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"liblmx error {code}: {msg}")
        self.code = code


class lmx_ray(ctypes.Structure):
    _fields_ = [("o", ctypes.c_float * 3), ("tmin", ctypes.c_float),
                ("d", ctypes.c_float * 3), ("tmax", ctypes.c_float)]


class lmx_hit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("b0", ctypes.c_float), ("b1", ctypes.c_float),
                ("id", ctypes.c_uint32)]


class lmx_build_opts(ctypes.Structure):
    _fields_ = [("isect", ctypes.c_int32), ("nthreads", ctypes.c_int32)]


class lmx_accel_info(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("num_tris", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("isect", ctypes.c_int32),
                ("node_bytes", ctypes.c_uint64), ("tri_bytes", ctypes.c_uint64),
                ("stack_overflow", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class lmx_material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("R", ctypes.c_float * 3), ("exponent", ctypes.c_float),
                ("Le", ctypes.c_float * 3), ("roughness", ctypes.c_float), ("eta", ctypes.c_float * 3),
                ("k", ctypes.c_float * 3), ("reserved", ctypes.c_float)]


class lmx_camera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("vx", ctypes.c_float * 3), ("vy", ctypes.c_float * 3),
                ("vz", ctypes.c_float * 3), ("fov_deg", ctypes.c_float), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32)]


class lmx_scene_desc(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("nvert", ctypes.c_uint32),
                ("tris", ctypes.c_void_p), ("ntri", ctypes.c_uint32),
                ("normals", ctypes.c_void_p), ("tri_mat", ctypes.c_void_p),
                ("mats", ctypes.c_void_p), ("nmat", ctypes.c_uint32),
                ("light_prims", ctypes.c_void_p), ("nlight", ctypes.c_uint32),
                ("camera", lmx_camera)]


class lmx_render_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("min_num_vertices", ctypes.c_int32),
                ("rr_depth", ctypes.c_int32), ("rr_prob", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("num_samples", ctypes.c_uint64),
                ("raster_mode", ctypes.c_int32), ("queue_capacity", ctypes.c_uint32),
                ("uniform_table", ctypes.c_void_p), ("table_width", ctypes.c_uint32),
                ("sample_out", ctypes.c_void_p), ("rr_mode", ctypes.c_int32), ("mis", ctypes.c_int32),
                ("render_time", ctypes.c_double), ("shade_order", ctypes.c_int32),
                ("ray_out_cap", ctypes.c_uint32), ("ray_out", ctypes.c_void_p)]


class lmx_render_stats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("iterations", ctypes.c_uint32),
                ("queue_capacity", ctypes.c_uint32), ("ms", ctypes.c_double), ("samples", ctypes.c_uint64),
                ("dropped", ctypes.c_uint64)]


class lmx_vcm_params(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("num_iteration_pass", ctypes.c_uint32),
                ("num_photon_trace_samples", ctypes.c_uint32), ("initial_radius", ctypes.c_float),
                ("alpha", ctypes.c_float), ("photon_seed", ctypes.c_uint64), ("photon_table", ctypes.c_void_p),
                ("photon_table_width", ctypes.c_uint32)]


class lmx_pm_params(ctypes.Structure):
    _fields_ = [("num_photon_trace_samples", ctypes.c_uint32), ("finalgather", ctypes.c_int32),
                ("radius", ctypes.c_float), ("photon_seed", ctypes.c_uint64), ("photon_table", ctypes.c_void_p),
                ("photon_table_width", ctypes.c_uint32)]


class lmx_ppm_params(ctypes.Structure):
    _fields_ = [("num_iteration_pass", ctypes.c_uint32), ("num_photon_trace_samples", ctypes.c_uint32),
                ("initial_radius", ctypes.c_float), ("alpha", ctypes.c_float), ("photon_seed", ctypes.c_uint64),
                ("photon_table", ctypes.c_void_p), ("photon_table_width", ctypes.c_uint32)]


# Every symbol include/lmx.h declares, with its ctypes signature.
class lmx_guide_opts(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("batch", ctypes.c_uint32), ("steps_per_pass", ctypes.c_uint32),
                ("wclamp", ctypes.c_float), ("seed", ctypes.c_uint64)]


class lmx_guide_stats(ctypes.Structure):
    _fields_ = [("records", ctypes.c_uint32), ("records_nonzero", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("loss", ctypes.c_float),
                ("render_ms", ctypes.c_double), ("train_ms", ctypes.c_double)]


SIGNATURES = {
    "lmx_last_error": (ctypes.c_char_p, []),
    "lmx_accel_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.POINTER(lmx_build_opts), ctypes.POINTER(ctypes.c_void_p)]),
    "lmx_accel_free": (None, [ctypes.c_void_p]),
    "lmx_accel_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_accel_info)]),
    "lmx_accel_upload": (ctypes.c_int, [ctypes.c_void_p]),
    "lmx_accel_intersect": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                           ctypes.c_void_p, ctypes.c_void_p]),
    "lmx_accel_occluded": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "lmx_accel_intersect_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "lmx_accel_intersect_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_void_p, ctypes.c_int32]),
    "lmx_scene_create": (ctypes.c_int, [ctypes.POINTER(lmx_scene_desc), ctypes.POINTER(ctypes.c_void_p)]),
    "lmx_scene_free": (None, [ctypes.c_void_p]),
    "lmx_render_pt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_bdpt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_vcm_bdpt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_vcm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params),
                                      ctypes.POINTER(lmx_vcm_params), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_pm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params),
                                     ctypes.POINTER(lmx_pm_params), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_ppm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params),
                                      ctypes.POINTER(lmx_ppm_params), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_sppm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params),
                                       ctypes.POINTER(lmx_ppm_params), ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_ltdirect": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_raycast": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.POINTER(lmx_render_stats)]),
    "lmx_film_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                     ctypes.c_void_p]),
    "lmx_image_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "lmx_render_ptdirect": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_render_ptmis": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_render_params), ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.POINTER(lmx_render_stats)]),
    "lmx_guide_default_opts": (None, [ctypes.POINTER(lmx_guide_opts)]),
    "lmx_guide_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(lmx_guide_opts),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "lmx_guide_free": (None, [ctypes.c_void_p]),
    "lmx_guide_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "lmx_render_guided": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(lmx_render_params),
                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(lmx_guide_stats), ctypes.POINTER(lmx_render_stats)]),
    "lmx_guide_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "lmx_guide_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "lmx_guide_mix_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "lmx_accel_export": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "lmx_fmath_eval": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "lmx_accel_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.c_int32]),
    "lmx_accel_count_occluded": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                                ctypes.c_int32]),
}


def _load():
    # torch loads its bundled HIP runtime under the name libamdhip64.so with
    # soname libamdhip64.so.7.  Importing torch first makes liblmx's
    # dependency on libamdhip64.so.7 resolve to that same, already loaded
    # runtime; loading liblmx first would bring in /opt/rocm's copy as well,
    # and two HIP runtimes in one process do not share devices.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python __graft_entry__.py build` "
            "(the HIP extension is required; there is no CPU fallback)")
    so = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(so, name)
        fn.restype = res
        fn.argtypes = args
    return so


lib = _load()


def check(rc):
    if rc != 0:
        raise LmxError(rc, lib.lmx_last_error().decode(errors="replace"))
    return rc
