"""ctypes harness over libmrt_hip.so (include/mrt.h, include/mrt_host.h).

The host-side mirror of the reference's render back-end interface:
``Renderer`` wraps init_shader / convert_scene_build / compute_frame_gpu /
get_result / clean_shaders (src/gpu/*.c) as mrt_init / mrt_upload_scene /
mrt_render / mrt_readback_* / mrt_destroy, and ``HostScene`` wraps the scene
front-end (the .rt reader and the procedural generators).  Every failing
call raises ``MrtError`` carrying the library's status code and message.
There is no fallback: if the shared library is missing, importing this
module raises.
"""
from __future__ import annotations

import ctypes as C  # noqa: N812
import os
from pathlib import Path

import numpy as np

_PKG_DIR = Path(__file__).resolve().parents[2]          # 42-minirt_amd/
LIB_PATH = Path(os.environ.get("MRT_HIP_LIB", _PKG_DIR / "libmrt_hip.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"libmrt_hip.so not found at {LIB_PATH}; build it first "
        "(python -c 'import __graft_entry__ as g; g.build()')")

_lib = C.CDLL(str(LIB_PATH))

# ---- status codes (mrt.h) ---------------------------------------------------
MRT_OK = 0
MRT_E_INVALID = -1
MRT_E_NOMEM = -2
MRT_E_HIP = -3
MRT_E_STATE = -4
MRT_E_UNSUPPORTED = -5
MRT_E_LIMIT = -6
MRT_E_NODEVICE = -7

PRIM_NONE = 0xFFFFFFFF
PRIM_TRIANGLE = 0x00000000
PRIM_SPHERE = 0x40000000
PRIM_PLANE = 0x80000000
PRIM_QUADRIC = 0xC0000000
PRIM_KIND_MASK = 0xC0000000

LIGHT_POINT = 0
LIGHT_DIRECTIONAL = 1
SCENE_TRIANGLE_ONE_SIDED = 0x1
SCENE_LIGHT_EXACT = 0x2
EST_REFERENCE = 0
EST_MIS = 1

OPT_SEED = 1
OPT_TRACE_TOP = 11
OPT_TRACE_REFILL = 12
OPT_ESTIMATOR = 2
OPT_STATS = 3
OPT_PATH_BUDGET = 4
OPT_BLOCK = 5
OPT_PHOTON_K = 6
OPT_PHOTON_NORM = 7
OPT_BUILDER = 8
OPT_LANES = 10
BUILDER_HOST = 0
BUILDER_DEVICE = 1
STATS_COUNT = 1
STATS_TIME = 2

AOV_PRIM_ID = 0
AOV_T = 1
AOV_NORMAL = 2
AOV_NODES = 3
AOV_PRIMS = 4


class MrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mrt error {code}: {msg}")
        self.code = code


# ---- structs ---------------------------------------------------------------
class Material(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("metallic", C.c_float),
                ("roughness", C.c_float), ("ior", C.c_float),
                ("transmission", C.c_float), ("ao", C.c_float),
                ("emission_strength", C.c_float),
                ("emission_color", C.c_float * 3)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("vec", C.c_float * 3),
                ("brightness", C.c_float), ("color", C.c_float * 3),
                ("radius", C.c_float), ("shadow_samples", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3),
                ("fov_deg", C.c_float), ("defocus_angle_deg", C.c_float),
                ("focus_dist", C.c_float)]


_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)


QUADRIC_CYLINDER = 0
QUADRIC_CONE = 1
QUADRIC_HYPERBOLOID = 2


class Quadric(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pos", C.c_float * 3),
                ("axis", C.c_float * 3), ("height", C.c_float),
                ("radius", C.c_float), ("a", C.c_float), ("b", C.c_float),
                ("c", C.c_float), ("shape", C.c_float), ("mat", C.c_uint32),
                ("mat_top", C.c_uint32), ("mat_bot", C.c_uint32)]


assert C.sizeof(Quadric) == 64


TEX_IMAGE = 0
TEX_CHECKER_LOCAL = 1
TEX_CHECKER_GLOBAL = 2
TEX_BLACK = 3
NO_TEX = 0xFFFFFFFF


class Texture(C.Structure):
    _fields_ = [("type", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("texel0", C.c_uint32),
                ("scale", C.c_float), ("c1", C.c_float * 3),
                ("c2", C.c_float * 3), ("pad", C.c_uint32)]


class MaterialTex(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "color", "metallic", "roughness", "transmission", "ao",
        "emission_strength", "emission_color", "normal")] + \
        [("normal_intensity", C.c_float), ("scale", C.c_float)]


assert C.sizeof(Texture) == 48 and C.sizeof(MaterialTex) == 40


class Scene(C.Structure):
    _fields_ = [("n_triangles", C.c_uint32), ("tri_pos", _fp),
                ("tri_nrm", _fp), ("tri_mat", _up),
                ("n_spheres", C.c_uint32), ("sph", _fp), ("sph_mat", _up),
                ("n_planes", C.c_uint32), ("pln", _fp), ("pln_mat", _up),
                ("n_materials", C.c_uint32),
                ("materials", C.POINTER(Material)),
                ("n_lights", C.c_uint32), ("lights", C.POINTER(Light)),
                ("camera", Camera), ("ambient_ratio", C.c_float),
                ("ambient_color", C.c_float * 3), ("ior_all", C.c_float),
                ("flags", C.c_uint32), ("n_quadrics", C.c_uint32),
                ("quadrics", C.POINTER(Quadric)),
                ("n_textures", C.c_uint32), ("textures", C.POINTER(Texture)),
                ("n_texels", C.c_uint32), ("texels", C.POINTER(C.c_uint8)),
                ("material_tex", C.POINTER(MaterialTex)),
                ("tri_uv", _fp)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "paths", "rays_closest", "rays_shadow", "nodes_closest",
        "prims_closest", "nodes_shadow", "prims_shadow", "launches_trace",
        "launches_shadow", "launches_shade", "nonfinite")] + \
        [(n, C.c_double) for n in (
            "ms_raygen", "ms_trace", "ms_shade", "ms_shadow", "ms_accum",
            "ms_total", "ms_caustic")] + \
        [(n, C.c_uint64) for n in ("gather_queries", "gather_photons",
                                   "gather_cells", "gather_lookups")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchTuning(C.Structure):
    _fields_ = [("path_budget", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("use_graph", C.c_uint32),
                ("trace_refill", C.c_uint32), ("ms_per_mpath", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BvhNode(C.Structure):
    _fields_ = [("lo0", C.c_float * 3), ("hi0", C.c_float * 3),
                ("lo1", C.c_float * 3), ("hi1", C.c_float * 3),
                ("child", C.c_int32 * 2), ("pad", C.c_uint32 * 2)]


class Prim(C.Structure):
    _fields_ = [("a", C.c_float * 4), ("b", C.c_float * 4),
                ("c", C.c_float * 4)]


class Bvh(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("nodes", C.POINTER(BvhNode)),
                ("n_prims", C.c_uint32), ("prims", C.POINTER(Prim)),
                ("shade_nrm", _fp), ("depth", C.c_uint32),
                ("n_leaves", C.c_uint32), ("sah_cost", C.c_double)]


class Bvh4(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("nodes", C.c_void_p),
                ("depth", C.c_uint32)]


class HdriTables(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("pdf_joint", _fp),
                ("marg_inv", _fp), ("cond_inv", _fp),
                ("column_average", C.c_double)]


assert C.sizeof(Material) == 48 and C.sizeof(Light) == 40
assert C.sizeof(BvhNode) == 64 and C.sizeof(Prim) == 48


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_ctx_p = C.c_void_p
_hs_p = C.c_void_p
_vp = C.c_void_p

mrt_init = _sig("mrt_init", C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                C.POINTER(_ctx_p))
mrt_upload_scene = _sig("mrt_upload_scene", C.c_int, _ctx_p, C.POINTER(Scene))
mrt_set_camera = _sig("mrt_set_camera", C.c_int, _ctx_p, C.POINTER(Camera))
mrt_upload_hdri = _sig("mrt_upload_hdri", C.c_int, _ctx_p, _fp, C.c_uint32,
                       C.c_uint32, C.c_float)
mrt_build_photons = _sig("mrt_build_photons", C.c_int, _ctx_p, C.c_uint32,
                         C.c_float)
mrt_get_photons = _sig("mrt_get_photons", C.c_int, _ctx_p, C.c_uint32, _fp,
                       _fp, _up, _up)
mrt_caustic_at = _sig("mrt_caustic_at", C.c_int, _ctx_p, C.c_uint32, _fp, _fp)
mrt_diff_set_materials = _sig("mrt_diff_set_materials", C.c_int, _ctx_p,
                              C.c_uint32, _fp, _fp, _fp)
mrt_diff_set_hdri = _sig("mrt_diff_set_hdri", C.c_int, _ctx_p, _fp)
mrt_adjoint = _sig("mrt_adjoint", C.c_int, _ctx_p, C.c_uint32, C.c_uint32,
                   C.c_uint32, _fp, _fp, _fp)
mrt_render = _sig("mrt_render", C.c_int, _ctx_p, C.c_uint32, C.c_uint32,
                  C.c_uint32)
mrt_readback_add = _sig("mrt_readback_add", C.c_int, _ctx_p,
                        C.POINTER(C.c_double))
mrt_readback_f32 = _sig("mrt_readback_f32", C.c_int, _ctx_p, _fp)
mrt_clear = _sig("mrt_clear", C.c_int, _ctx_p)
mrt_render_aov = _sig("mrt_render_aov", C.c_int, _ctx_p, C.c_int, C.c_uint32,
                      _vp)
mrt_trace_rays = _sig("mrt_trace_rays", C.c_int, _ctx_p, C.c_uint32, _fp, _fp,
                      _up, _fp)
mrt_trace_rays_noalloc = _sig("mrt_trace_rays_noalloc", C.c_int, _ctx_p,
                              C.c_uint32, _fp, _fp, _up, _fp)
TRACE_NOALLOC_MAX = 262144
mrt_set_option = _sig("mrt_set_option", C.c_int, _ctx_p, C.c_int, C.c_double)
mrt_set_seed = _sig("mrt_set_seed", C.c_int, _ctx_p, C.c_uint64)
mrt_reserve = _sig("mrt_reserve", C.c_int, _ctx_p, C.c_uint32)
mrt_get_stats = _sig("mrt_get_stats", C.c_int, _ctx_p, C.POINTER(Stats))
mrt_reset_stats = _sig("mrt_reset_stats", C.c_int, _ctx_p)
mrt_synchronize = _sig("mrt_synchronize", C.c_int, _ctx_p)
mrt_destroy = _sig("mrt_destroy", None, _ctx_p)
mrt_last_error = _sig("mrt_last_error", C.c_char_p, _ctx_p)
mrt_version = _sig("mrt_version", C.c_char_p)

_bt_p = C.c_void_p
mrt_batch_create = _sig("mrt_batch_create", C.c_int, C.c_int, C.POINTER(_bt_p))
mrt_batch_add_scene = _sig("mrt_batch_add_scene", C.c_int, _bt_p,
                           C.POINTER(Scene), C.c_uint32, C.c_uint32,
                           C.c_uint32, C.c_uint64, _up)
mrt_batch_commit = _sig("mrt_batch_commit", C.c_int, _bt_p)
mrt_batch_autotune = _sig("mrt_batch_autotune", C.c_int, _bt_p, C.c_uint32,
                          C.POINTER(BatchTuning))
mrt_batch_set_tuning = _sig("mrt_batch_set_tuning", C.c_int, _bt_p,
                            C.POINTER(BatchTuning))
mrt_batch_get_tuning = _sig("mrt_batch_get_tuning", C.c_int, _bt_p,
                            C.POINTER(BatchTuning))
mrt_batch_render = _sig("mrt_batch_render", C.c_int, _bt_p, C.c_uint32)
mrt_batch_readback = _sig("mrt_batch_readback", C.c_int, _bt_p, C.c_uint32,
                          _fp)
mrt_batch_clear = _sig("mrt_batch_clear", C.c_int, _bt_p)
mrt_batch_set_option = _sig("mrt_batch_set_option", C.c_int, _bt_p, C.c_int,
                            C.c_double)
mrt_batch_get_stats = _sig("mrt_batch_get_stats", C.c_int, _bt_p,
                           C.POINTER(Stats))
mrt_batch_destroy = _sig("mrt_batch_destroy", None, _bt_p)
mrt_batch_last_error = _sig("mrt_batch_last_error", C.c_char_p, _bt_p)

mrt_bvh_build = _sig("mrt_bvh_build", C.c_int, C.POINTER(Scene), C.c_uint32,
                     C.POINTER(Bvh))
mrt_bvh_free = _sig("mrt_bvh_free", None, C.POINTER(Bvh))
mrt_bvh4_build = _sig("mrt_bvh4_build", C.c_int, C.POINTER(Bvh),
                      C.POINTER(Bvh4))
mrt_bvh4_free = _sig("mrt_bvh4_free", None, C.POINTER(Bvh4))
mrt_hs_new = _sig("mrt_hs_new", _hs_p)
mrt_hs_free = _sig("mrt_hs_free", None, _hs_p)
mrt_hs_add_material = _sig("mrt_hs_add_material", C.c_int, _hs_p,
                           C.POINTER(Material), C.c_char_p)
mrt_hs_find_material = _sig("mrt_hs_find_material", C.c_int, _hs_p,
                            C.c_char_p)
mrt_hs_add_sphere = _sig("mrt_hs_add_sphere", C.c_int, _hs_p, _fp, C.c_float,
                         C.c_uint32)
mrt_hs_add_quadric = _sig("mrt_hs_add_quadric", C.c_int, _hs_p,
                          C.POINTER(Quadric))
mrt_hs_add_texture = _sig("mrt_hs_add_texture", C.c_int, _hs_p,
                          C.POINTER(Texture), C.POINTER(C.c_uint8), C.c_char_p)
mrt_hs_set_material_tex = _sig("mrt_hs_set_material_tex", C.c_int, _hs_p,
                               C.c_uint32, C.POINTER(MaterialTex))
mrt_hs_set_triangle_uv = _sig("mrt_hs_set_triangle_uv", C.c_int, _hs_p,
                              C.c_uint32, _fp)
mrt_get_scene_bvh = _sig("mrt_get_scene_bvh", C.c_int, C.c_void_p,
                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                         C.POINTER(C.c_uint32), C.POINTER(C.c_int),
                         C.c_void_p, C.c_void_p, _fp)
mrt_get_hdri_tables = _sig("mrt_get_hdri_tables", C.c_int, C.c_void_p,
                           *([_fp] * 6))
mrt_decode_image = _sig("mrt_decode_image", C.c_int, C.c_char_p, C.c_size_t,
                        C.POINTER(C.POINTER(C.c_uint8)),
                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32))
mrt_hs_add_plane = _sig("mrt_hs_add_plane", C.c_int, _hs_p, _fp, _fp,
                        C.c_uint32)
mrt_hs_add_triangle = _sig("mrt_hs_add_triangle", C.c_int, _hs_p, _fp, _fp,
                           C.c_uint32)
mrt_hs_add_light = _sig("mrt_hs_add_light", C.c_int, _hs_p, C.POINTER(Light))
mrt_hs_set_camera = _sig("mrt_hs_set_camera", None, _hs_p, C.POINTER(Camera))
mrt_hs_set_ambient = _sig("mrt_hs_set_ambient", None, _hs_p, C.c_float, _fp)
mrt_hs_set_resolution = _sig("mrt_hs_set_resolution", None, _hs_p,
                             C.c_uint32, C.c_uint32)
mrt_hs_get_resolution = _sig("mrt_hs_get_resolution", None, _hs_p,
                             _up, _up)
mrt_hs_set_flags = _sig("mrt_hs_set_flags", None, _hs_p, C.c_uint32)
mrt_hs_set_hdri = _sig("mrt_hs_set_hdri", C.c_int, _hs_p, _fp, C.c_uint32,
                       C.c_uint32, C.c_float)
mrt_hs_get_hdri = _sig("mrt_hs_get_hdri", C.c_int, _hs_p, C.POINTER(_fp),
                       _up, _up, C.POINTER(C.c_float))
mrt_hs_view = _sig("mrt_hs_view", C.POINTER(Scene), _hs_p)
mrt_gen_plain_scene = _sig("mrt_gen_plain_scene", C.c_int, C.c_uint64,
                           C.c_uint32, C.c_uint32, C.POINTER(_hs_p))
mrt_gen_mesh_scene = _sig("mrt_gen_mesh_scene", C.c_int, C.c_uint64,
                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                          C.c_float, C.POINTER(_hs_p))
mrt_gen_caustic_scene = _sig("mrt_gen_caustic_scene", C.c_int, C.c_uint64,
                             C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.c_uint32, C.POINTER(_hs_p))
mrt_gen_icosphere = _sig("mrt_gen_icosphere", C.c_int, _hs_p, _fp, C.c_float,
                         C.c_uint32, C.c_float, C.c_uint64, C.c_int,
                         C.c_uint32)
mrt_gen_icosphere_freq = _sig("mrt_gen_icosphere_freq", C.c_int, _hs_p, _fp,
                              C.c_float, C.c_uint32, C.c_float, C.c_uint64,
                              C.c_int, C.c_uint32)
mrt_gen_hdri_sky = _sig("mrt_gen_hdri_sky", C.c_int, C.c_uint64, C.c_uint32,
                        C.c_uint32, C.c_float, C.c_float, _fp)
mrt_hdri_build_tables = _sig("mrt_hdri_build_tables", C.c_int, _fp,
                             C.c_uint32, C.c_uint32, C.c_float,
                             C.POINTER(HdriTables))
mrt_hdri_free_tables = _sig("mrt_hdri_free_tables", None,
                            C.POINTER(HdriTables))
mrt_parse_rt_file = _sig("mrt_parse_rt_file", C.c_int, C.c_char_p,
                         C.POINTER(_hs_p), C.c_char_p, C.c_size_t)
mrt_parse_rt_string = _sig("mrt_parse_rt_string", C.c_int, C.c_char_p,
                           C.POINTER(_hs_p), C.c_char_p, C.c_size_t)
mrt_decode_hdr = _sig("mrt_decode_hdr", C.c_int, C.c_char_p, C.c_size_t,
                      C.POINTER(_fp), _up, _up, C.POINTER(C.c_float),
                      C.POINTER(C.c_float))
mrt_free = _sig("mrt_free", None, C.c_void_p)
mrt_hs_add_obj_text = _sig("mrt_hs_add_obj_text", C.c_int, _hs_p, C.c_char_p,
                           _fp, _fp, C.c_float, C.c_uint32)
mrt_heat_map_add = _sig("mrt_heat_map_add", None, C.POINTER(C.c_int32),
                        C.c_uint32, C.c_uint32, C.POINTER(C.c_double))
mrt_export_ppm = _sig("mrt_export_ppm", C.c_int, C.c_char_p,
                      C.POINTER(C.c_double), C.c_uint32, C.c_uint32,
                      C.c_uint32, C.c_double)
mrt_export_rts = _sig("mrt_export_rts", C.c_int, C.c_char_p,
                      C.POINTER(C.c_double), C.c_uint32, C.c_uint32,
                      C.c_uint32)
mrt_load_rts = _sig("mrt_load_rts", C.c_int, C.c_char_p,
                    C.POINTER(C.c_double), C.c_uint32, C.c_uint32, _up)


class FrameState(C.Structure):
    """mrt_frame: one render_frame loop's state (include/mrt_host.h)."""
    _fields_ = [("ctx", C.c_void_p), ("w", C.c_uint32), ("h", C.c_uint32),
                ("sample", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32),
                ("samples_per_tick", C.c_uint32),
                ("float_render", C.POINTER(C.c_double))]


mrt_frame_tick = _sig("mrt_frame_tick", C.c_int, C.POINTER(FrameState))


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_up)


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype).copy()


def _check(rc, ctx=None):
    if rc < 0:
        msg = mrt_last_error(ctx).decode() if ctx else ""
        raise MrtError(rc, msg)
    return rc


# ---- host scenes -------------------------------------------------------------
class HostScene:
    """Owns an mrt_host_scene (the library frees it on close)."""

    def __init__(self, handle):
        if not handle:
            raise MrtError(MRT_E_NOMEM, "no scene")
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) \
            else handle

    @classmethod
    def empty(cls):
        return cls(mrt_hs_new())

    @classmethod
    def plain(cls, seed, width=640, height=360):
        h = _hs_p()
        _check(mrt_gen_plain_scene(seed, width, height, C.byref(h)))
        return cls(h)

    @classmethod
    def mesh(cls, seed, width=1920, height=1080, subdiv=8, floor_cells=256,
             amp=0.15):
        h = _hs_p()
        _check(mrt_gen_mesh_scene(seed, width, height, subdiv, floor_cells,
                                  amp, C.byref(h)))
        return cls(h)

    @classmethod
    def caustic(cls, seed, width=1920, height=1080, subdiv=8, floor_cells=256,
                water_cells=512):
        h = _hs_p()
        _check(mrt_gen_caustic_scene(seed, width, height, subdiv, floor_cells,
                                     water_cells, C.byref(h)))
        return cls(h)

    @classmethod
    def parse(cls, text):
        h = _hs_p()
        err = C.create_string_buffer(512)
        rc = mrt_parse_rt_string(text.encode(), C.byref(h), err, 512)
        if rc < 0:
            raise MrtError(rc, err.value.decode())
        return cls(h)

    @classmethod
    def parse_file(cls, path):
        h = _hs_p()
        err = C.create_string_buffer(512)
        rc = mrt_parse_rt_file(str(path).encode(), C.byref(h), err, 512)
        if rc < 0:
            raise MrtError(rc, err.value.decode())
        return cls(h)

    def close(self):
        if self._h:
            mrt_hs_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def view(self):
        return mrt_hs_view(self._h)

    @property
    def resolution(self):
        w, h = C.c_uint32(), C.c_uint32()
        mrt_hs_get_resolution(self._h, C.byref(w), C.byref(h))
        return w.value, h.value

    def set_resolution(self, w, h):
        mrt_hs_set_resolution(self._h, w, h)

    def set_flags(self, flags):
        mrt_hs_set_flags(self._h, flags)

    def set_camera(self, pos, direction, fov=60.0, defocus=0.0, focus=1.0):
        cam = Camera((C.c_float * 3)(*pos), (C.c_float * 3)(*direction),
                     fov, defocus, focus)
        mrt_hs_set_camera(self._h, C.byref(cam))

    def set_ambient(self, ratio, rgb255):
        _, p = _f32(rgb255)
        a = np.asarray(rgb255, dtype=np.float32)
        mrt_hs_set_ambient(self._h, ratio, a.ctypes.data_as(_fp))

    def add_material(self, color=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.0,
                     ior=0.0, transmission=0.0, ao=1.0, emission=0.0,
                     emission_color=(0.0, 0.0, 0.0), name=None):
        m = Material((C.c_float * 3)(*color), metallic, roughness, ior,
                     transmission, ao, emission,
                     (C.c_float * 3)(*emission_color))
        return _check(mrt_hs_add_material(
            self._h, C.byref(m), name.encode() if name else None))

    def add_quadric(self, kind, pos, axis, height, mat, radius=0.0, a=0.0,
                    b=0.0, c=0.0, shape=0.0, mat_top=None, mat_bot=None):
        """A cylinder, cone or hyperboloid (mrt_quadric); cap materials
        default to mat."""
        q = Quadric()
        q.kind = kind
        q.pos[:] = [float(x) for x in pos]
        q.axis[:] = [float(x) for x in axis]
        q.height, q.radius = height, radius
        q.a, q.b, q.c, q.shape = a, b, c, shape
        q.mat = mat
        q.mat_top = mat if mat_top is None else mat_top
        q.mat_bot = mat if mat_bot is None else mat_bot
        return _check(mrt_hs_add_quadric(self._h, C.byref(q)))

    def add_cylinder(self, centre, axis, diameter, height, mat, mat_top=None,
                     mat_bot=None):
        """The .rt "cy" element: centre, axis, diameter, height."""
        return self.add_quadric(QUADRIC_CYLINDER, centre, axis, height, mat,
                                radius=0.5 * diameter, mat_top=mat_top,
                                mat_bot=mat_bot)

    def add_cone(self, apex, axis, diameter, height, mat, mat_top=None):
        """The .rt "co" element: apex, axis, base diameter, height."""
        return self.add_quadric(QUADRIC_CONE, apex, axis, height, mat,
                                radius=0.5 * diameter, mat_top=mat_top)

    def add_hyperboloid(self, centre, axis, height, a, b, c, shape, mat):
        """The .rt "hb" element: centre, axis, height, a, b, c, shape."""
        return self.add_quadric(QUADRIC_HYPERBOLOID, centre, axis, height, mat,
                                a=a, b=b, c=c, shape=shape)

    def add_image_texture(self, rgba, name=None):
        """An IMAGE texture from an (h, w, 4) uint8 array (row 0 first)."""
        a = np.ascontiguousarray(rgba, dtype=np.uint8)
        t = Texture()
        t.type = TEX_IMAGE
        t.height, t.width = a.shape[0], a.shape[1]
        return _check(mrt_hs_add_texture(
            self._h, C.byref(t), a.ctypes.data_as(C.POINTER(C.c_uint8)),
            name.encode() if name else None))

    def add_checker_texture(self, kind, scale, c1, c2, name=None):
        """A CHECKER_LOCAL / CHECKER_GLOBAL texture (colours 0..255)."""
        t = Texture()
        t.type = kind
        t.scale = scale
        t.c1[:] = [float(x) for x in c1]
        t.c2[:] = [float(x) for x in c2]
        return _check(mrt_hs_add_texture(self._h, C.byref(t), None,
                                         name.encode() if name else None))

    def set_material_tex(self, mat, normal_intensity=1.0, scale=1.0,
                         **slots):
        """Texture slots of a material: color, metallic, roughness,
        transmission, ao, emission_strength, emission_color, normal."""
        mt = MaterialTex()
        for f, _ in MaterialTex._fields_[:8]:
            setattr(mt, f, slots.pop(f, NO_TEX))
        if slots:
            raise TypeError(f"unknown texture slots {sorted(slots)}")
        mt.normal_intensity = normal_intensity
        mt.scale = scale
        _check(mrt_hs_set_material_tex(self._h, mat, C.byref(mt)))

    def set_triangle_uv(self, tri, uv):
        a = np.ascontiguousarray(uv, dtype=np.float32).reshape(6)
        _check(mrt_hs_set_triangle_uv(self._h, tri, a.ctypes.data_as(_fp)))

    def add_sphere(self, centre, radius, mat):
        c = np.asarray(centre, dtype=np.float32)
        return _check(mrt_hs_add_sphere(self._h, c.ctypes.data_as(_fp),
                                        radius, mat))

    def add_plane(self, point, normal, mat):
        p = np.asarray(point, dtype=np.float32)
        n = np.asarray(normal, dtype=np.float32)
        return _check(mrt_hs_add_plane(self._h, p.ctypes.data_as(_fp),
                                       n.ctypes.data_as(_fp), mat))

    def add_triangle(self, v, mat, n=None):
        va = np.asarray(v, dtype=np.float32).reshape(9)
        na = None if n is None else np.asarray(n, dtype=np.float32).reshape(9)
        return _check(mrt_hs_add_triangle(
            self._h, va.ctypes.data_as(_fp),
            na.ctypes.data_as(_fp) if na is not None else None, mat))

    def add_light(self, pos, brightness=1.0, color=(255, 255, 255),
                  radius=0.0, samples=0, kind=LIGHT_POINT):
        L = Light(kind, (C.c_float * 3)(*pos), brightness,
                  (C.c_float * 3)(*color), radius, samples)
        return _check(mrt_hs_add_light(self._h, C.byref(L)))

    def add_icosphere(self, centre, radius, subdiv, amp, seed, mix=True,
                      mat=0, freq=None):
        """A displaced icosphere: 20 * 4^subdiv triangles, or with freq
        the geodesic sphere of 20 * freq^2."""
        c = np.asarray(centre, dtype=np.float32)
        if freq is not None:
            return _check(mrt_gen_icosphere_freq(self._h, c.ctypes.data_as(_fp),
                                                 radius, freq, amp, seed,
                                                 int(mix), mat))
        return _check(mrt_gen_icosphere(self._h, c.ctypes.data_as(_fp),
                                        radius, subdiv, amp, seed,
                                        int(mix), mat))

    def add_obj_text(self, text, pos, direction, scale, mat):
        p = np.asarray(pos, np.float32)
        d = np.asarray(direction, np.float32)
        return _check(mrt_hs_add_obj_text(self._h, text.encode(),
                                          p.ctypes.data_as(_fp),
                                          d.ctypes.data_as(_fp), scale, mat))

    def hdri(self):
        """(rgb[h, w, 3], exposure) of the scene's HDRI, or None."""
        p = _fp()
        w, h, ex = C.c_uint32(0), C.c_uint32(0), C.c_float(0)
        if mrt_hs_get_hdri(self._h, C.byref(p), C.byref(w), C.byref(h),
                           C.byref(ex)) < 0:
            return None
        a = np.ctypeslib.as_array(p, shape=(h.value, w.value, 3)).copy()
        return a, ex.value

    def set_hdri(self, rgb, exposure=0.0):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w = rgb.shape[:2]
        _check(mrt_hs_set_hdri(self._h, rgb.ctypes.data_as(_fp), w, h,
                               exposure))

    def arrays(self):
        """numpy copies of the scene arrays (for the oracle and tests)."""
        s = self.view().contents
        return {
            "tri_pos": _arr(s.tri_pos, 9 * s.n_triangles, np.float32)
            .reshape(-1, 9),
            "tri_nrm": _arr(s.tri_nrm, 9 * s.n_triangles, np.float32)
            .reshape(-1, 9),
            "tri_mat": _arr(s.tri_mat, s.n_triangles, np.uint32),
            "sph": _arr(s.sph, 4 * s.n_spheres, np.float32).reshape(-1, 4),
            "sph_mat": _arr(s.sph_mat, s.n_spheres, np.uint32),
            "pln": _arr(s.pln, 6 * s.n_planes, np.float32).reshape(-1, 6),
            "pln_mat": _arr(s.pln_mat, s.n_planes, np.uint32),
            "quadrics": [
                {f: (list(getattr(s.quadrics[i], f))
                     if f in ("pos", "axis") else getattr(s.quadrics[i], f))
                 for f, _ in Quadric._fields_}
                for i in range(s.n_quadrics)],
        }


class SceneBvh:
    """A scene's BVH as read back from a Renderer: 4-wide nodes and records
    only (no binary tree)."""
    nodes = np.zeros(0, np.uint8)
    n_nodes = 0

    def node4_floats(self):
        return self.nodes4.view(np.float32).reshape(-1, 32)

    def node4_children(self):
        return self.nodes4.view(np.int32).reshape(-1, 32)[:, 24:28]

    def prim_words(self):
        return self.prims.view(np.uint32).reshape(-1, 12)


class BuiltBvh:
    """mrt_bvh_build result as numpy arrays (host, no GPU)."""

    def __init__(self, scene_view, max_leaf=4):
        b = Bvh()
        _check(mrt_bvh_build(scene_view, max_leaf, C.byref(b)))
        try:
            self.n_nodes = b.n_nodes
            self.n_prims = b.n_prims
            self.depth = b.depth
            self.n_leaves = b.n_leaves
            self.sah_cost = b.sah_cost
            self.nodes = (np.ctypeslib.as_array(
                C.cast(b.nodes, C.POINTER(C.c_uint8)),
                shape=(64 * b.n_nodes,)).copy() if b.n_nodes else
                np.zeros(0, np.uint8))
            self.prims = (np.ctypeslib.as_array(
                C.cast(b.prims, C.POINTER(C.c_uint8)),
                shape=(48 * b.n_prims,)).copy() if b.n_prims else
                np.zeros(0, np.uint8))
            self.shade_nrm = (np.ctypeslib.as_array(
                b.shade_nrm, shape=(9 * b.n_prims,)).copy() if b.n_prims else
                np.zeros(0, np.float32))
            b4 = Bvh4()
            _check(mrt_bvh4_build(C.byref(b), C.byref(b4)))
            try:
                self.n_nodes4 = b4.n_nodes
                self.depth4 = b4.depth
                self.nodes4 = (np.ctypeslib.as_array(
                    C.cast(b4.nodes, C.POINTER(C.c_uint8)),
                    shape=(128 * b4.n_nodes,)).copy() if b4.n_nodes else
                    np.zeros(0, np.uint8))
            finally:
                mrt_bvh4_free(C.byref(b4))
        finally:
            mrt_bvh_free(C.byref(b))

    def node_floats(self):
        return self.nodes.view(np.float32).reshape(-1, 16)

    def node_children(self):
        return self.nodes.view(np.int32).reshape(-1, 16)[:, 12:14]

    def node4_floats(self):
        return self.nodes4.view(np.float32).reshape(-1, 32)

    def node4_children(self):
        return self.nodes4.view(np.int32).reshape(-1, 32)[:, 24:28]

    def prim_words(self):
        return self.prims.view(np.uint32).reshape(-1, 12)


def hdri_tables(rgb, exposure=0.0):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    t = HdriTables()
    _check(mrt_hdri_build_tables(rgb.ctypes.data_as(_fp), w, h, exposure,
                                 C.byref(t)))
    try:
        out = (np.ctypeslib.as_array(t.pdf_joint, shape=(h * w,)).copy()
               .reshape(h, w),
               np.ctypeslib.as_array(t.marg_inv, shape=(h,)).copy(),
               np.ctypeslib.as_array(t.cond_inv, shape=(h * w,)).copy()
               .reshape(h, w), t.column_average)
    finally:
        mrt_hdri_free_tables(C.byref(t))
    return out


def hdri_sky(seed, w, h, sun_radius_deg=0.5, sun_radiance=5e4):
    out = np.zeros((h, w, 3), dtype=np.float32)
    _check(mrt_gen_hdri_sky(seed, w, h, sun_radius_deg, sun_radiance,
                            out.ctypes.data_as(_fp)))
    return out


# ---- the device renderer -------------------------------------------------------
class Renderer:
    """One mrt_ctx: the replacement of the reference's GL render back-end."""

    def __init__(self, width, height, device=0):
        ctx = _ctx_p()
        rc = mrt_init(device, width, height, C.byref(ctx))
        if rc < 0:
            raise MrtError(rc, "mrt_init failed (no gfx950 device?)")
        self._ctx = ctx
        self.width, self.height = width, height

    def close(self):
        if getattr(self, "_ctx", None):
            mrt_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        return _check(rc, self._ctx)

    def upload(self, scene):
        view = scene.view() if isinstance(scene, HostScene) else scene
        self._ck(mrt_upload_scene(self._ctx, view))

    def set_option(self, opt, value):
        self._ck(mrt_set_option(self._ctx, opt, float(value)))

    def set_seed(self, seed):
        """The full 64-bit RNG key (mrt_set_seed)."""
        self._ck(mrt_set_seed(self._ctx, int(seed) & (2 ** 64 - 1)))

    def upload_hdri(self, rgb, exposure=0.0):
        a = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w = a.shape[:2]
        self._ck(mrt_upload_hdri(self._ctx, a.ctypes.data_as(_fp), w, h,
                                 exposure))
        self._env_wh = (w, h)

    def scene_bvh(self):
        """The uploaded scene's BVH (SceneBvh: 4-wide nodes, primitive and
        shading records as the kernels read them)."""
        nn, npr, dep, dev = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int()
        self._ck(mrt_get_scene_bvh(self._ctx, C.byref(nn), C.byref(npr),
                                   C.byref(dep), C.byref(dev), None, None,
                                   None))
        b = SceneBvh()
        b.n_nodes4, b.n_prims, b.depth4 = nn.value, npr.value, dep.value
        b.on_device = bool(dev.value)
        b.nodes4 = np.zeros(128 * b.n_nodes4, np.uint8)
        b.prims = np.zeros(48 * b.n_prims, np.uint8)
        b.shade_nrm = np.zeros(9 * b.n_prims, np.float32)
        self._ck(mrt_get_scene_bvh(self._ctx, None, None, None, None,
                                   b.nodes4.ctypes.data, b.prims.ctypes.data,
                                   b.shade_nrm.ctypes.data_as(_fp)))
        return b

    def hdri_tables(self):
        """The device's HDRI tables: (pdf_joint, marg_inv, cond_inv,
        mis_marg, mis_cond, mis_pcell)."""
        w, h = self._env_wh
        out = [np.zeros((h, w), np.float32), np.zeros(h, np.float32),
               np.zeros((h, w), np.float32), np.zeros(h, np.float32),
               np.zeros((h - 1, w), np.float32),
               np.zeros((h - 1, w - 1), np.float32)]
        self._ck(mrt_get_hdri_tables(self._ctx, *[a.ctypes.data_as(_fp)
                                                  for a in out]))
        return tuple(out)

    def build_photons(self, n_per, radius, k=None, norm=None):
        """Emit, sort and grid the caustic photon map (--caustic=N;R;NORM)."""
        if k is not None:
            self.set_option(OPT_PHOTON_K, k)
        if norm is not None:
            self.set_option(OPT_PHOTON_NORM, norm)
        self._ck(mrt_build_photons(self._ctx, n_per, radius))

    def photon_count(self):
        n = C.c_uint32(0)
        self._ck(mrt_get_photons(self._ctx, 0, None, None, None, C.byref(n)))
        return n.value

    def photons(self):
        """Stored photons in map order: (pos[n,3], power[n,3], ids[n])."""
        n = C.c_uint32(0)
        self._ck(mrt_get_photons(self._ctx, 0, None, None, None, C.byref(n)))
        pos = np.zeros((n.value, 3), np.float32)
        pw = np.zeros((n.value, 3), np.float32)
        ids = np.zeros(n.value, np.uint32)
        self._ck(mrt_get_photons(self._ctx, n.value, pos.ctypes.data_as(_fp),
                                 pw.ctypes.data_as(_fp),
                                 ids.ctypes.data_as(_up), C.byref(n)))
        return pos, pw, ids

    def caustic_at(self, pts):
        p, pp = _f32(pts)
        n = p.size // 3
        out = np.zeros((n, 3), np.float32)
        self._ck(mrt_caustic_at(self._ctx, n, pp, out.ctypes.data_as(_fp)))
        return out

    def diff_set_materials(self, albedo=None, roughness=None, emission=None):
        n = None
        arrs = []
        for a, k in ((albedo, 3), (roughness, 1), (emission, 3)):
            if a is None:
                arrs.append((None, None))
                continue
            a = np.ascontiguousarray(a, np.float32).reshape(-1, k)
            n = len(a)
            arrs.append((a, a.ctypes.data_as(_fp)))
        if n is None:
            return
        self._ck(mrt_diff_set_materials(self._ctx, n, arrs[0][1], arrs[1][1],
                                        arrs[2][1]))

    def diff_set_hdri(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        self._ck(mrt_diff_set_hdri(self._ctx, a.ctypes.data_as(_fp)))

    def adjoint(self, first_sample, n_samples, max_depth, d_image, n_materials,
                hdri_shape=None):
        """dL/dparams for dL/dI of the mean image: (g_mat[n, 7] = albedo rgb,
        roughness, emission rgb; g_hdri[h, w, 3] or None)."""
        di = np.ascontiguousarray(d_image, np.float32)
        gm = np.zeros((n_materials, 7), np.float32)
        gt = None
        if hdri_shape is not None:
            gt = np.zeros((hdri_shape[0], hdri_shape[1], 3), np.float32)
        self._ck(mrt_adjoint(self._ctx, first_sample, n_samples, max_depth,
                             di.ctypes.data_as(_fp), gm.ctypes.data_as(_fp),
                             gt.ctypes.data_as(_fp) if gt is not None else None))
        return gm, gt

    def set_camera(self, pos, direction, fov=60.0, defocus=0.0, focus=1.0):
        """mrt_set_camera: a new viewpoint for the uploaded scene."""
        cam = Camera((C.c_float * 3)(*pos), (C.c_float * 3)(*direction), fov,
                     defocus, focus)
        self._ck(mrt_set_camera(self._ctx, C.byref(cam)))

    def reserve(self, n_samples):
        """Allocate the queues of renders of n_samples per call now."""
        self._ck(mrt_reserve(self._ctx, n_samples))

    def render(self, first_sample, n_samples, max_depth):
        self._ck(mrt_render(self._ctx, first_sample, n_samples, max_depth))

    def synchronize(self):
        self._ck(mrt_synchronize(self._ctx))

    def clear(self):
        self._ck(mrt_clear(self._ctx))

    def readback(self):
        out = np.zeros((self.height, self.width, 3), dtype=np.float32)
        self._ck(mrt_readback_f32(self._ctx, out.ctypes.data_as(_fp)))
        return out

    def readback_add(self, float_render):
        """get_result: adds the samples rendered since the last call into
        the double buffer and zeroes the device sums."""
        assert float_render.dtype == np.float64 and float_render.flags.c_contiguous
        self._ck(mrt_readback_add(
            self._ctx, float_render.ctypes.data_as(C.POINTER(C.c_double))))

    def frame(self, float_render, spp, max_bounces, samples_per_tick=1,
              sample=0):
        """A render_frame loop over this context (mrt_frame_tick)."""
        return Frame(self, float_render, spp, max_bounces, samples_per_tick,
                     sample)

    def aov(self, kind, sample=0):
        n = self.width * self.height
        if kind == AOV_T:
            out = np.zeros(n, np.float32)
        elif kind == AOV_NORMAL:
            out = np.zeros((n, 3), np.float32)
        else:
            out = np.zeros(n, np.uint32)
        self._ck(mrt_render_aov(self._ctx, kind, sample,
                                out.ctypes.data_as(C.c_void_p)))
        return out

    def trace_rays(self, orig, dirs):
        o, po = _f32(orig)
        d, pd = _f32(dirs)
        n = o.size // 3
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        self._ck(mrt_trace_rays(self._ctx, n, po, pd,
                                prim.ctypes.data_as(_up),
                                t.ctypes.data_as(_fp)))
        return prim, t

    def trace_rays_noalloc(self, orig, dirs):
        """mrt_trace_rays through the library's fixed staging array (no
        device allocation; at most TRACE_NOALLOC_MAX rays)."""
        o, po = _f32(orig)
        d, pd = _f32(dirs)
        n = o.size // 3
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        self._ck(mrt_trace_rays_noalloc(self._ctx, n, po, pd,
                                        prim.ctypes.data_as(_up),
                                        t.ctypes.data_as(_fp)))
        return prim, t

    def stats(self):
        s = Stats()
        self._ck(mrt_get_stats(self._ctx, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._ck(mrt_reset_stats(self._ctx))


class Frame:
    """render_frame ticks (src/rendering/loop.c:36-63) into a host-owned
    double buffer through mrt_frame_tick."""

    def __init__(self, renderer, float_render, spp, max_bounces,
                 samples_per_tick=1, sample=0):
        assert float_render.dtype == np.float64
        assert float_render.flags.c_contiguous
        self._r = renderer
        self.buf = float_render
        self.state = FrameState(renderer._ctx, renderer.width, renderer.height,
                                sample, spp, max_bounces, samples_per_tick,
                                float_render.ctypes.data_as(
                                    C.POINTER(C.c_double)))

    @property
    def sample(self):
        return self.state.sample

    def tick(self):
        """Samples added by this tick (0 once spp is reached)."""
        return self._r._ck(mrt_frame_tick(C.byref(self.state)))


class Batch:
    """Many scenes rendered through one set of wavefront queues."""

    def __init__(self, device=0):
        h = _bt_p()
        rc = mrt_batch_create(device, C.byref(h))
        if rc < 0:
            raise MrtError(rc, "mrt_batch_create failed (no gfx950 device?)")
        self._b = h
        self.sizes = []

    def close(self):
        if getattr(self, "_b", None):
            mrt_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        if rc < 0:
            raise MrtError(rc, mrt_batch_last_error(self._b).decode())
        return rc

    def add(self, scene, width, height, spp, seed):
        view = scene.view() if isinstance(scene, HostScene) else scene
        sid = C.c_uint32(0)
        self._ck(mrt_batch_add_scene(self._b, view, width, height, spp, seed,
                                     C.byref(sid)))
        self.sizes.append((width, height))
        return sid.value

    def commit(self):
        self._ck(mrt_batch_commit(self._b))

    def autotune(self, max_depth):
        t = BatchTuning()
        self._ck(mrt_batch_autotune(self._b, max_depth, C.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = BatchTuning()
        self._ck(mrt_batch_get_tuning(self._b, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = BatchTuning()
        self._ck(mrt_batch_get_tuning(self._b, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._ck(mrt_batch_set_tuning(self._b, C.byref(t)))

    def set_option(self, opt, value):
        self._ck(mrt_batch_set_option(self._b, opt, float(value)))

    def render(self, max_depth):
        self._ck(mrt_batch_render(self._b, max_depth))

    def readback(self, sid):
        w, h = self.sizes[sid]
        out = np.zeros((h, w, 3), np.float32)
        self._ck(mrt_batch_readback(self._b, sid, out.ctypes.data_as(_fp)))
        return out

    def clear(self):
        self._ck(mrt_batch_clear(self._b))

    def stats(self):
        s = Stats()
        self._ck(mrt_batch_get_stats(self._b, C.byref(s)))
        return s.as_dict()


def decode_hdr(data):
    """Radiance .hdr bytes -> (rgb[h, w, 3] float32, exposure, gamma)."""
    p = _fp()
    w, h = C.c_uint32(0), C.c_uint32(0)
    ex, ga = C.c_float(0), C.c_float(0)
    _check(mrt_decode_hdr(data, len(data), C.byref(p), C.byref(w), C.byref(h),
                          C.byref(ex), C.byref(ga)))
    try:
        a = np.ctypeslib.as_array(p, shape=(h.value, w.value, 3)).copy()
    finally:
        mrt_free(p)
    return a, ex.value, ga.value


def decode_image(data):
    """BMP / PNG bytes -> (h, w, 4) uint8, rows as the reference's readers
    order them."""
    p = C.POINTER(C.c_uint8)()
    w, h = C.c_uint32(), C.c_uint32()
    _check(mrt_decode_image(data, len(data), C.byref(p), C.byref(w),
                            C.byref(h)))
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        mrt_free(C.cast(p, C.c_void_p))


def heat_map_add(depth, float_render):
    """set_heat_map_color: float_render (h, w, 3 doubles) += ramp colour of
    depth / max(depth) where depth >= 0 (in place)."""
    d = np.ascontiguousarray(depth, dtype=np.int32)
    a = float_render
    assert a.dtype == np.float64 and a.flags.c_contiguous
    h, w = a.shape[:2]
    assert d.size == w * h
    mrt_heat_map_add(d.ctypes.data_as(C.POINTER(C.c_int32)), w, h,
                     a.ctypes.data_as(C.POINTER(C.c_double)))
    return a


def export_ppm(path, float_render, samples, gamma=1.0):
    a = np.ascontiguousarray(float_render, dtype=np.float64)
    h, w = a.shape[:2]
    _check(mrt_export_ppm(str(path).encode(),
                          a.ctypes.data_as(C.POINTER(C.c_double)), w, h,
                          samples, gamma))


def export_rts(path, float_render, samples):
    a = np.ascontiguousarray(float_render, dtype=np.float64)
    h, w = a.shape[:2]
    _check(mrt_export_rts(str(path).encode(),
                          a.ctypes.data_as(C.POINTER(C.c_double)), w, h,
                          samples))


def load_rts(path, w, h):
    a = np.zeros((h, w, 3), np.float64)
    s = C.c_uint32()
    _check(mrt_load_rts(str(path).encode(),
                        a.ctypes.data_as(C.POINTER(C.c_double)), w, h,
                        C.byref(s)))
    return a, s.value


def version():
    return mrt_version().decode()
