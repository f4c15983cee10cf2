"""ORACLE (test infrastructure only): ctypes access to liboracle.so.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
as the checker or the timed CPU baseline -- never by the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

TRAV_BRUTE = 0
TRAV_REFERENCE = 1
TRAV_ORDERED = 2


class OrcBvh(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("n_nodes", C.c_uint32),
                ("prims", C.c_void_p), ("n_prims", C.c_uint32),
                ("nodes4", C.c_void_p), ("n_nodes4", C.c_uint32)]


class OrcHdri(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgb", C.c_void_p),
                ("exposure", C.c_float), ("pdf_joint", C.c_void_p),
                ("marg_inv", C.c_void_p), ("cond_inv", C.c_void_p),
                ("mis_marg", C.c_void_p), ("mis_cond", C.c_void_p),
                ("mis_pcell", C.c_void_p)]


P_ALBEDO, P_ROUGHNESS, P_EMISSION, P_TEXEL = 0, 1, 2, 3


class OrcParam(C.Structure):
    _fields_ = [("kind", C.c_int), ("mat", C.c_uint32), ("chan", C.c_uint32),
                ("texel", C.c_uint32)]


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("first_sample", C.c_uint32), ("n_samples", C.c_uint32),
                ("max_depth", C.c_uint32), ("seed", C.c_uint64),
                ("traversal", C.c_int), ("estimator", C.c_int),
                ("threads", C.c_int), ("row_begin", C.c_uint32),
                ("row_end", C.c_uint32)]


class OrcCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "rays_closest", "rays_shadow", "nodes_closest", "prims_closest",
        "nodes_shadow", "prims_shadow")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(str(LIB))
        vp, u32, fp = C.c_void_p, C.c_uint32, C.POINTER(C.c_float)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_uniform4.argtypes = [C.c_uint64, u32, u32, u32, u32, u32, fp]
        L.orc_viewport.argtypes = [vp, u32, u32, C.POINTER(C.c_double)]
        L.orc_raygen_f32.argtypes = [vp, u32, u32, u32, C.c_uint64, fp, fp]
        for name in ("orc_trace_rays_f32", "orc_trace_rays_f64"):
            getattr(L, name).argtypes = [vp, vp, C.c_int, u32, fp, fp,
                                         C.POINTER(C.c_uint32), vp,
                                         C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32), C.c_int]
        for name in ("orc_hit_info_f32", "orc_hit_info_f64"):
            getattr(L, name).argtypes = [vp, vp, C.c_int, u32, fp, fp,
                                         C.POINTER(C.c_uint32), vp, vp,
                                         C.POINTER(C.c_uint8),
                                         C.POINTER(C.c_uint32), fp]
        L.orc_occluded_f32.argtypes = [vp, vp, u32, fp, fp, fp,
                                       C.POINTER(C.c_uint8),
                                       C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]
        for name in ("orc_render_f32", "orc_render_f64"):
            getattr(L, name).argtypes = [vp, vp, vp, vp, C.POINTER(OrcParams),
                                         C.POINTER(C.c_double),
                                         C.POINTER(OrcCounters)]
            getattr(L, name).restype = C.c_int
        for name in ("orc_emit_photons_f32", "orc_emit_photons_f64"):
            getattr(L, name).argtypes = [vp, vp, vp, u32, C.c_uint64, fp, fp,
                                         C.POINTER(C.c_uint32)]
            getattr(L, name).restype = u32
        L.orc_caustic_f32.argtypes = [vp, u32, fp, fp]
        for name in ("orc_render_grad_f32", "orc_render_grad_f64"):
            getattr(L, name).argtypes = [vp, vp, vp, C.POINTER(OrcParams),
                                         C.POINTER(OrcParam),
                                         C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
            getattr(L, name).restype = C.c_int
        L.orc_hdri_tables.argtypes = [fp, u32, u32, C.c_float, fp, fp, fp,
                                      C.POINTER(C.c_double)]
        L.orc_hdri_tables.restype = C.c_int
        dp = C.POINTER(C.c_double)
        L.orc_sample_ggx_f64.argtypes = [dp, C.c_double, C.c_float,
                                         C.c_float, dp]
        L.orc_sample_cos_f64.argtypes = [dp, C.c_float, C.c_float, dp]
        L.orc_unit_sphere_f64.argtypes = [fp, dp]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def uniform4(seed, pixel, sample, bounce, block, stream=0):
    o = np.zeros(4, np.float32)
    lib().orc_uniform4(seed, pixel, sample, bounce, block, stream, _fp(o))
    return o


def viewport(camera_ptr, w, h):
    o = (C.c_double * 19)()
    lib().orc_viewport(camera_ptr, w, h, o)
    return np.array(o[:])


def raygen_f32(scene_view, w, h, sample, seed):
    cam = C.addressof(scene_view.contents.camera)
    o = np.zeros((w * h, 3), np.float32)
    d = np.zeros((w * h, 3), np.float32)
    lib().orc_raygen_f32(cam, w, h, sample, seed, _fp(o), _fp(d))
    return o, d


def make_bvh(built, wide=True):
    """OrcBvh over a mrt_hip.BuiltBvh (keeps a reference to its arrays); the
    ordered traversal walks the 4-wide nodes unless wide is False."""
    b = OrcBvh(built.nodes.ctypes.data, built.n_nodes,
               built.prims.ctypes.data, built.n_prims,
               built.nodes4.ctypes.data if wide else None,
               built.n_nodes4 if wide else 0)
    b._keep = built
    return b


def trace_rays(scene_view, bvh, mode, orig, dirs, f64=False, threads=0):
    o = np.ascontiguousarray(orig, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = o.size // 3
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float64 if f64 else np.float32)
    nodes = np.zeros(n, np.uint32)
    prims = np.zeros(n, np.uint32)
    fn = lib().orc_trace_rays_f64 if f64 else lib().orc_trace_rays_f32
    fn(C.cast(scene_view, C.c_void_p),
       C.byref(bvh) if bvh is not None else None, mode, n, _fp(o), _fp(d),
       _up(prim), t.ctypes.data, _up(nodes), _up(prims), threads)
    return prim, t, nodes, prims


def hit_info(scene_view, bvh, mode, orig, dirs, f64=False, albedo=False):
    """Closest hits with normal (facing the ray), front flag and material
    (and the textured colour when albedo)."""
    o = np.ascontiguousarray(orig, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = o.size // 3
    ft = np.float64 if f64 else np.float32
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, ft)
    nrm = np.zeros((n, 3), ft)
    front = np.zeros(n, np.uint8)
    mat = np.zeros(n, np.uint32)
    alb = np.zeros((n, 3), np.float32)
    fn = lib().orc_hit_info_f64 if f64 else lib().orc_hit_info_f32
    fn(C.cast(scene_view, C.c_void_p),
       C.byref(bvh) if bvh is not None else None, mode, n, _fp(o), _fp(d),
       _up(prim), t.ctypes.data, nrm.ctypes.data,
       front.ctypes.data_as(C.POINTER(C.c_uint8)), _up(mat), _fp(alb))
    if albedo:
        return prim, t, nrm, front.astype(bool), mat, alb
    return prim, t, nrm, front.astype(bool), mat


def occluded_f32(scene_view, bvh, orig, dirs, tmax):
    o = np.ascontiguousarray(orig, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    tm = np.ascontiguousarray(tmax, np.float32)
    n = o.size // 3
    occ = np.zeros(n, np.uint8)
    nodes = np.zeros(n, np.uint32)
    prims = np.zeros(n, np.uint32)
    lib().orc_occluded_f32(C.cast(scene_view, C.c_void_p), C.byref(bvh), n,
                           _fp(o), _fp(d), _fp(tm),
                           occ.ctypes.data_as(C.POINTER(C.c_uint8)),
                           _up(nodes), _up(prims))
    return occ, nodes, prims


class OrcPhotonMap(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("pow", C.c_void_p), ("n", C.c_uint32),
                ("k", C.c_uint32), ("radius", C.c_float), ("norm", C.c_float)]


def photon_map(pos, pw, k=15, radius=0.01, norm=1.0):
    pos = np.ascontiguousarray(pos, np.float32)
    pw = np.ascontiguousarray(pw, np.float32)
    pm = OrcPhotonMap(pos.ctypes.data, pw.ctypes.data, len(pos), k, radius,
                      norm)
    pm._keep = (pos, pw)
    return pm


def emit_photons(scene_view, bvh, n_per, seed, hdri=None, n_emitters=1,
                 f64=False):
    cap = max(1, n_emitters) * n_per
    pos = np.zeros((cap, 3), np.float32)
    pw = np.zeros((cap, 3), np.float32)
    gid = np.zeros(cap, np.uint32)
    fn = lib().orc_emit_photons_f64 if f64 else lib().orc_emit_photons_f32
    n = fn(C.cast(scene_view, C.c_void_p), C.byref(bvh),
           C.byref(hdri) if hdri is not None else None, n_per, seed,
           _fp(pos), _fp(pw), gid.ctypes.data_as(C.POINTER(C.c_uint32)))
    return pos[:n].copy(), pw[:n].copy(), gid[:n].copy()


def caustic(pm, pts):
    p = np.ascontiguousarray(pts, np.float32)
    out = np.zeros_like(p)
    lib().orc_caustic_f32(C.byref(pm), len(p), _fp(p), _fp(out))
    return out


def render(scene_view, bvh, width, height, first_sample, n_samples,
           max_depth, seed, f64=False, traversal=TRAV_ORDERED,
           estimator=0, threads=0, rows=None, hdri=None, photons=None):
    acc = np.zeros((height, width, 3), np.float64)
    p = OrcParams(width, height, first_sample, n_samples, max_depth, seed,
                  traversal, estimator, threads,
                  rows[0] if rows else 0, rows[1] if rows else 0)
    cnt = OrcCounters()
    fn = lib().orc_render_f64 if f64 else lib().orc_render_f32
    rc = fn(C.cast(scene_view, C.c_void_p),
            C.byref(bvh) if bvh is not None else None,
            C.byref(hdri) if hdri is not None else None,
            C.byref(photons) if photons is not None else None, C.byref(p),
            acc.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cnt))
    if rc != 0:
        raise RuntimeError("oracle render failed")
    return acc, cnt.as_dict()


def render_grad(scene_view, bvh, width, height, first_sample, n_samples,
                max_depth, seed, kind, mat=0, chan=0, texel=0, f64=False,
                hdri=None, threads=0):
    """Forward-mode derivative of the sum image with respect to one
    parameter: returns (image sums, derivative sums)."""
    img = np.zeros((height, width, 3), np.float64)
    dimg = np.zeros((height, width, 3), np.float64)
    p = OrcParams(width, height, first_sample, n_samples, max_depth, seed,
                  TRAV_ORDERED, 1, threads, 0, 0)
    q = OrcParam(kind, mat, chan, texel)
    fn = lib().orc_render_grad_f64 if f64 else lib().orc_render_grad_f32
    rc = fn(C.cast(scene_view, C.c_void_p),
            C.byref(bvh) if bvh is not None else None,
            C.byref(hdri) if hdri is not None else None, C.byref(p),
            C.byref(q), img.ctypes.data_as(C.POINTER(C.c_double)),
            dimg.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("oracle render_grad failed")
    return img, dimg


def hdri_tables(rgb, exposure=0.0):
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    pdf = np.zeros((h, w), np.float32)
    marg = np.zeros(h, np.float32)
    cond = np.zeros((h, w), np.float32)
    col = C.c_double()
    rc = lib().orc_hdri_tables(_fp(rgb), w, h, exposure, _fp(pdf), _fp(marg),
                               _fp(cond), C.byref(col))
    if rc != 0:
        raise RuntimeError("oracle hdri tables failed")
    return pdf, marg, cond, col.value


def sample_ggx(n, rough, u1, u2):
    nn = (C.c_double * 3)(*n)
    o = (C.c_double * 3)()
    lib().orc_sample_ggx_f64(nn, rough, u1, u2, o)
    return np.array(o[:])


def sample_cos(n, u1, u2):
    nn = (C.c_double * 3)(*n)
    o = (C.c_double * 3)()
    lib().orc_sample_cos_f64(nn, u1, u2, o)
    return np.array(o[:])


def unit_sphere(u4):
    u = np.ascontiguousarray(u4, np.float32)
    o = (C.c_double * 3)()
    lib().orc_unit_sphere_f64(_fp(u), o)
    return np.array(o[:])


def hdri_mis_tables(rgb, exposure=0.0):
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    marg = np.zeros(h, np.float32)
    cond = np.zeros((h - 1, w), np.float32)
    pcell = np.zeros((h - 1, w - 1), np.float32)
    tot = C.c_double()
    L = lib()
    L.orc_hdri_mis_tables.argtypes = [C.POINTER(C.c_float), C.c_uint32,
                                      C.c_uint32, C.c_float] + \
        [C.POINTER(C.c_float)] * 3 + [C.POINTER(C.c_double)]
    if L.orc_hdri_mis_tables(_fp(rgb), w, h, exposure, _fp(marg), _fp(cond),
                             _fp(pcell), C.byref(tot)) != 0:
        raise RuntimeError("oracle MIS tables failed")
    return marg, cond, pcell, tot.value


def make_hdri(rgb, exposure=0.0):
    """OrcHdri with the oracle's own reference and MIS tables."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    pdf, marg, cond, _ = hdri_tables(rgb, exposure)
    mm, mc, mp, _ = hdri_mis_tables(rgb, exposure)
    hd = OrcHdri(w, h, rgb.ctypes.data, exposure, pdf.ctypes.data,
                 marg.ctypes.data, cond.ctypes.data, mm.ctypes.data,
                 mc.ctypes.data, mp.ctypes.data)
    hd._keep = (rgb, pdf, marg, cond, mm, mc, mp)
    return hd


def refract(d, n, front, ior, ior_all=1.0, color=(1.0, 1.0, 1.0), f64=True):
    """Transmission branch of a smooth dielectric for one direction d at a
    surface whose normal n faces it: (new direction, power, totally
    reflected)."""
    L = lib()
    fn = L.orc_refract_f64 if f64 else L.orc_refract_f32
    dp = C.POINTER(C.c_double)
    fn.argtypes = [dp, dp, C.c_int, C.c_float, C.c_float,
                   C.POINTER(C.c_float), dp, dp]
    fn.restype = C.c_int
    dd = (C.c_double * 3)(*map(float, d))
    nn = (C.c_double * 3)(*map(float, n))
    col = (C.c_float * 3)(*map(float, color))
    out = (C.c_double * 3)()
    pw = (C.c_double * 3)()
    tir = fn(dd, nn, int(front), float(ior), float(ior_all), col, out, pw)
    return np.array(out[:]), np.array(pw[:]), bool(tir)
