"""ctypes binding of libillum_hiprt.so (the C ABI in include/illum_hiprt.h).

Host objects (scenes, BVHs) come back as numpy arrays; device buffers are
passed as torch CUDA tensors, whose memory and streams are plumbing only --
every trace runs in the library's own gfx950 kernels.  Loading fails loudly
when the shared library is missing; there is no fallback path.

Import torch before this module in processes that use both, so that the one
HIP runtime of the process is torch's.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libillum_hiprt.so")

BVH2, BVH4 = 0, 1
TRI_EXACT, TRI_WOOP = 0, 1
SPLIT_OM, SPLIT_SM, SPLIT_SAH = 0, 1, 2

TRACERS = ("while-while", "persistent-while-while", "persistent-speculative-while-while",
           "dynamic-while-while")
# "<tracer>-alpha": the alpha-aware variant (hit only where texture alpha > 0.5)
TRACER_ALPHA = 0x100

# numpy view of the reference's 56-byte bbvh_node (rt/cpu/bvh.h:42-52)
BBVH_NODE = np.dtype([("box_l", np.float32, (2, 3)), ("box_r", np.float32, (2, 3)),
                      ("link_l", np.int32), ("link_r", np.int32)])
assert BBVH_NODE.itemsize == 56


class HiprtError(RuntimeError):
    pass


class _BvhParams(ctypes.Structure):
    _fields_ = [("split", c_int32), ("max_tris", c_int32), ("planes", c_int32), ("esc", c_int32)]


class _Rays(ctypes.Structure):
    _fields_ = [(n, c_void_p) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class _Hits(ctypes.Structure):
    _fields_ = [("t", c_void_p), ("u", c_void_p), ("v", c_void_p), ("id", c_void_p)]


class _TracerOpts(ctypes.Structure):
    _fields_ = [("refill_threshold", c_int32), ("blocks", c_int32), ("d_fetch_stats", c_void_p),
                ("stack_config", c_int32), ("work_chunks", c_int32), ("spill_limit", c_int32),
                ("park_leaves", c_int32)]


class _BatchEntry(ctypes.Structure):
    _fields_ = [("scene", c_void_p), ("rays", _Rays), ("hits", _Hits), ("occluded", c_void_p), ("n", c_uint32)]


class _TuneResult(ctypes.Structure):
    _fields_ = [("tracer", c_int32), ("refill_threshold", c_int32), ("blocks", c_int32), ("bvh_format", c_int32),
                ("mrays_per_s", c_float), ("candidates", c_int32), ("stack_config", c_int32),
                ("work_chunks", c_int32), ("park_leaves", c_int32)]


class _PtTuneResult(ctypes.Structure):
    _fields_ = [("tracer", c_int32), ("stack_config", c_int32), ("samples", c_int32), ("bvh_format", c_int32),
                ("park_leaves", c_int32), ("mrays_per_s", c_float), ("candidates", c_int32)]


class _PtParams(ctypes.Structure):
    _fields_ = [("max_depth", c_int32), ("rr_start", c_int32), ("rr_survive", c_float), ("seed", c_uint64),
                ("tracer", c_int32), ("batched", c_int32), ("stack_config", c_int32), ("work_chunks", c_int32),
                ("refill_threshold", c_int32), ("blocks", c_int32), ("ray_order", c_int32), ("samples", c_int32),
                ("emitters_reflect", c_int32), ("park_leaves", c_int32), ("pixel_centre", c_int32),
                ("regen_pool", c_int32), ("refill_coherent", c_int32)]


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HiprtError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "hiprt_last_error": (c_char_p, []),
            "hiprt_version": (c_char_p, []),
            "hiprt_camera_rays_host": (c_int, [POINTER(c_float), POINTER(c_void_p)]),
            "hiprt_scene_generate": (c_int, [c_int, POINTER(c_void_p), POINTER(c_float)]),
            "hiprt_scene_random": (c_int, [c_uint32, c_float, c_uint64, POINTER(c_void_p)]),
            "hiprt_scene_from_arrays": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p,
                                                POINTER(c_void_p)]),
            "hiprt_scene_load_obj": (c_int, [c_char_p, POINTER(c_void_p)]),
            "hiprt_scene_add_obj": (c_int, [POINTER(c_void_p), c_char_p, c_char_p, POINTER(c_float)]),
            "hiprt_scene_write_obj": (c_int, [c_void_p, c_char_p]),
            "hiprt_scene_info": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
            "hiprt_scene_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
            "hiprt_scene_materials": (c_int, [c_void_p, c_void_p, POINTER(c_uint32)]),
            "hiprt_scene_set_texcoords": (c_int, [c_void_p, c_void_p, c_uint32]),
            "hiprt_scene_get_texcoords": (c_int, [c_void_p, c_void_p, POINTER(c_uint32)]),
            "hiprt_scene_set_normals": (c_int, [c_void_p, c_void_p, c_uint32]),
            "hiprt_scene_set_positions": (c_int, [c_void_p, c_void_p, c_uint32]),
            "hiprt_scene_material_name": (c_int, [c_void_p, c_uint32, c_char_p, c_uint32]),
            "hiprt_scene_set_material": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
            "hiprt_scene_get_normals": (c_int, [c_void_p, c_void_p, POINTER(c_uint32)]),
            "hiprt_scene_set_texture": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_uint32]),
            "hiprt_scene_free": (None, [c_void_p]),
            "hiprt_bvh_build": (c_int, [c_void_p, POINTER(_BvhParams), POINTER(c_void_p)]),
            "hiprt_bvh_info": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                       POINTER(c_uint32)]),
            "hiprt_bvh_copy": (c_int, [c_void_p, c_void_p, c_void_p]),
            "hiprt_bvh_validate": (c_int, [c_void_p, c_void_p, c_int32]),
            "hiprt_bvh_free": (None, [c_void_p]),
            "hiprt_create": (c_int, [c_int32, POINTER(c_void_p)]),
            "hiprt_destroy": (c_int, [c_void_p]),
            "hiprt_scene_commit": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, POINTER(c_void_p)]),
            "hiprt_scene_upload": (c_int, [c_void_p, c_void_p, c_uint32, c_int32, c_void_p, c_void_p, c_uint32,
                                           c_int32, c_int32, POINTER(c_void_p)]),
            "hiprt_scene_layout": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int32),
                                           POINTER(c_uint32), POINTER(c_int32), POINTER(c_int32)]),
            "hiprt_scene_download": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
            "hiprt_scene_refit": (c_int, [c_void_p, c_void_p, c_void_p]),
            "hiprt_scene_build": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_void_p,
                                          c_uint32, c_int32, c_int32, POINTER(c_void_p), POINTER(c_float),
                                          c_void_p]),
            "hiprt_scene_destroy": (c_int, [c_void_p]),
            "hiprt_scene_stack_need": (c_int, [c_void_p, POINTER(c_uint32)]),
            "hiprt_tracer_id": (c_int, [c_char_p]),
            "hiprt_tracer_name": (c_char_p, [c_int]),
            "hiprt_trace_closest": (c_int, [c_void_p, POINTER(_Rays), c_uint32, POINTER(_Hits), c_int32,
                                            POINTER(_TracerOpts), c_void_p]),
            "hiprt_trace_any": (c_int, [c_void_p, POINTER(_Rays), c_uint32, c_void_p, POINTER(_Hits), c_int32,
                                        POINTER(_TracerOpts), c_void_p]),
            "hiprt_trace_closest_noalloc": (c_int, [c_void_p, POINTER(_Rays), c_uint32, POINTER(_Hits), c_int32,
                                                    POINTER(_TracerOpts), c_void_p]),
            "hiprt_camera_rays": (c_int, [POINTER(c_float), POINTER(_Rays), c_void_p]),
            "hiprt_incoherent_rays": (c_int, [POINTER(c_float), POINTER(c_float), c_float, c_float, c_float, c_uint64,
                                              POINTER(_Rays), c_uint32, c_void_p]),
            "hiprt_sort_keys_workspace": (ctypes.c_size_t, [c_uint32]),
            "hiprt_sort_keys": (c_int, [c_void_p, c_void_p, c_uint32, c_int, c_void_p, c_void_p]),
            "hiprt_bounce_cosine": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), c_uint32, c_uint64, c_int32,
                                            POINTER(_Rays), c_void_p]),
            "hiprt_pt_create": (c_int, [c_void_p, POINTER(c_float), POINTER(_PtParams), POINTER(c_void_p)]),
            "hiprt_pt_render": (c_int, [c_void_p, c_uint32, c_void_p]),
            "hiprt_pt_autotune": (c_int, [c_void_p, c_void_p, POINTER(c_float), POINTER(_PtParams), c_int32, c_int32,
                                          POINTER(_PtTuneResult)]),
            "hiprt_pt_ray_counts": (c_int, [c_void_p, POINTER(c_uint32), c_int32]),
            "hiprt_pt_image": (c_int, [c_void_p, c_void_p]),
            "hiprt_pt_memory": (c_int, [c_void_p, POINTER(c_uint64)]),
            "hiprt_pt_ray_totals": (c_int, [c_void_p, POINTER(c_uint64), c_int32]),
            "hiprt_pt_profile": (c_int, [c_void_p, c_uint32, c_void_p, POINTER(c_float), POINTER(c_uint64),
                                         POINTER(c_uint32), c_int32]),
            "hiprt_pt_reset": (c_int, [c_void_p, c_void_p]),
            "hiprt_pt_destroy": (c_int, [c_void_p]),
            "hiprt_adjoint_closest": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), c_uint32, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
            "hiprt_scene_counts": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
            "hiprt_trace_batch": (c_int, [POINTER(_BatchEntry), c_uint32, c_int32, c_int32, POINTER(_TracerOpts),
                                          c_void_p]),
            "hiprt_autotune": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(_Rays), c_uint32, c_int32,
                                       POINTER(_TuneResult), POINTER(c_void_p)]),
            "hiprt_add_hitpoint_albedo": (c_int, [c_void_p, POINTER(_Hits), c_uint32, c_void_p, c_void_p]),
            "hiprt_shadow_rays": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), c_uint32, POINTER(c_float),
                                          POINTER(_Rays), c_void_p]),
            "hiprt_integrate_light": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), POINTER(_Rays), c_void_p,
                                              c_uint32, POINTER(c_float), POINTER(c_float), c_void_p, c_void_p]),
            "hiprt_scene_set_lights": (c_int, [c_void_p, c_void_p, c_uint32, POINTER(c_uint32)]),
            "hiprt_sample_direct": (c_int, [c_void_p, c_int32, POINTER(_Rays), POINTER(_Hits), c_uint32, c_uint64,
                                            POINTER(_Rays), c_void_p, c_void_p, c_void_p, c_void_p]),
            "hiprt_integrate_dir_sample": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), POINTER(_Rays),
                                                   POINTER(_Hits), c_void_p, c_uint32, c_void_p, c_void_p]),
            "hiprt_integrate_light_sample": (c_int, [c_void_p, POINTER(_Rays), POINTER(_Hits), POINTER(_Rays),
                                                     c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names the C header declares (parsed from include/illum_hiprt.h)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "illum_hiprt.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(hiprt_[a-z0-9_]+)\s*\(", text)))


def check(status):
    if status != 0:
        raise HiprtError(lib().hiprt_last_error().decode())


def _ptr(a):
    return c_void_p(a.ctypes.data) if a is not None else None


class HostScene:
    """Host triangle scene (positions, index triples, materials)."""

    def __init__(self, handle, cam=None):
        self.h = c_void_p(handle) if not isinstance(handle, c_void_p) else handle
        self.cam = cam

    @classmethod
    def generate(cls, config):
        h = c_void_p()
        cam = (c_float * 15)()
        check(lib().hiprt_scene_generate(config, ctypes.byref(h), cam))
        return cls(h, np.array(cam, dtype=np.float32))

    @classmethod
    def random(cls, n, sigma=0.02, seed=1):
        h = c_void_p()
        check(lib().hiprt_scene_random(n, sigma, seed, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, pos, idx, mat=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        m = None if mat is None else np.ascontiguousarray(mat, dtype=np.uint32)
        h = c_void_p()
        check(lib().hiprt_scene_from_arrays(_ptr(pos), len(pos), _ptr(idx), len(idx), _ptr(m), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_obj(cls, path):
        h = c_void_p()
        check(lib().hiprt_scene_load_obj(str(path).encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def add_obj(cls, path, name="", trafo=None, scene=None):
        """hiprt_scene_add_obj: the OBJ file appended to `scene` (a new scene
        when None) under the 4x4 model matrix `trafo` (numpy layout, row
        major: trafo[row, column]; None = identity), its material names
        prefixed "<name>/" when a name is given.  Returns the scene."""
        h = scene.h if scene is not None else c_void_p()
        m = None
        if trafo is not None:
            m = np.ascontiguousarray(np.asarray(trafo, dtype=np.float32).reshape(4, 4).T)
        check(lib().hiprt_scene_add_obj(ctypes.byref(h), str(path).encode(), name.encode() if name else None,
                                        m.ctypes.data_as(POINTER(c_float)) if m is not None else None))
        return scene if scene is not None else cls(h)

    def write_obj(self, path):
        check(lib().hiprt_scene_write_obj(self.h, str(path).encode()))

    def arrays(self):
        nv, nt = c_uint32(), c_uint32()
        check(lib().hiprt_scene_info(self.h, ctypes.byref(nv), ctypes.byref(nt)))
        pos = np.empty((nv.value, 3), np.float32)
        idx = np.empty((nt.value, 3), np.uint32)
        mat = np.zeros(nt.value, np.uint32)
        check(lib().hiprt_scene_copy(self.h, _ptr(pos), _ptr(idx), _ptr(mat)))
        return pos, idx, mat

    def set_texcoords(self, tc):
        """Per-vertex texture coordinates (n_vert, 2) float32 (alpha-aware tracing)."""
        tc = np.ascontiguousarray(tc, dtype=np.float32).reshape(-1, 2)
        check(lib().hiprt_scene_set_texcoords(self.h, _ptr(tc), len(tc)))

    def texcoords(self):
        """Per-vertex texture coordinates (n_vert, 2) float32 (zeros when none)."""
        nv = c_uint32()
        check(lib().hiprt_scene_get_texcoords(self.h, None, ctypes.byref(nv)))
        tc = np.zeros((nv.value, 2), np.float32)
        check(lib().hiprt_scene_get_texcoords(self.h, _ptr(tc), ctypes.byref(nv)))
        return tc

    def material_names(self):
        """Names of the materials ("" where a scene's materials have none)."""
        n = len(self.materials())
        out = []
        for m in range(n):
            buf = ctypes.create_string_buffer(256)
            check(lib().hiprt_scene_material_name(self.h, m, buf, 256))
            out.append(buf.value.decode())
        return out

    def set_material(self, material, albedo=None, emissive=None):
        """Albedo and/or emission (rgb) of one material; None keeps a value."""
        a = None if albedo is None else np.ascontiguousarray(albedo, dtype=np.float32).reshape(3)
        e = None if emissive is None else np.ascontiguousarray(emissive, dtype=np.float32).reshape(3)
        check(lib().hiprt_scene_set_material(self.h, material, _ptr(a) if a is not None else None,
                                             _ptr(e) if e is not None else None))

    def set_positions(self, pos):
        """Vertex positions (n_vert, 3) float32, one per existing vertex."""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        check(lib().hiprt_scene_set_positions(self.h, _ptr(p), len(p)))

    def set_normals(self, normals):
        """Per-vertex shading normals (n_vert, 3) float32; None removes them."""
        if normals is None:
            check(lib().hiprt_scene_set_normals(self.h, None, 0))
            return
        n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        check(lib().hiprt_scene_set_normals(self.h, _ptr(n), len(n)))

    def normals(self):
        """Per-vertex shading normals (n_vert, 3) float32, or None when the
        scene has none."""
        nv = c_uint32()
        check(lib().hiprt_scene_get_normals(self.h, None, ctypes.byref(nv)))
        if nv.value == 0:
            return None
        n = np.zeros((nv.value, 3), np.float32)
        check(lib().hiprt_scene_get_normals(self.h, _ptr(n), ctypes.byref(nv)))
        return n

    def set_texture(self, material, rgba):
        """Attach an RGBA texture (h, w, 4) float32 as the albedo texture of a
        material (None detaches it); its alpha is the opacity alpha-aware
        tracers test."""
        if rgba is None:
            check(lib().hiprt_scene_set_texture(self.h, material, None, 0, 0))
            return
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        assert rgba.ndim == 3 and rgba.shape[2] == 4, rgba.shape
        check(lib().hiprt_scene_set_texture(self.h, material, _ptr(rgba), rgba.shape[1], rgba.shape[0]))

    def materials(self):
        """Material table (k, 8) float32: albedo rgb, 0, emission rgb, 0."""
        k = c_uint32()
        check(lib().hiprt_scene_materials(self.h, None, ctypes.byref(k)))
        table = np.zeros((k.value, 8), np.float32)
        if k.value:
            check(lib().hiprt_scene_materials(self.h, _ptr(table), ctypes.byref(k)))
        return table

    @property
    def n_tris(self):
        nv, nt = c_uint32(), c_uint32()
        check(lib().hiprt_scene_info(self.h, ctypes.byref(nv), ctypes.byref(nt)))
        return nt.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hiprt_scene_free(self.h)
            self.h = None


class Bvh:
    """Host BVH in the reference's node layout plus its leaf-slot index."""

    def __init__(self, scene, split=SPLIT_SAH, max_tris=4, planes=16, esc=True):
        p = _BvhParams(split, max_tris, planes, 1 if esc else 0)
        self.h = c_void_p()
        check(lib().hiprt_bvh_build(scene.h, ctypes.byref(p), ctypes.byref(self.h)))
        self.scene = scene
        self.max_tris = max_tris
        nn, ni, root, depth = c_uint32(), c_uint32(), c_uint32(), c_uint32()
        check(lib().hiprt_bvh_info(self.h, ctypes.byref(nn), ctypes.byref(ni), ctypes.byref(root),
                                   ctypes.byref(depth)))
        self.root, self.depth = root.value, depth.value
        self.nodes = np.zeros(nn.value, BBVH_NODE)
        self.index = np.zeros(ni.value, np.uint32)
        check(lib().hiprt_bvh_copy(self.h, _ptr(self.nodes), _ptr(self.index)))

    def validate(self):
        return lib().hiprt_bvh_validate(self.scene.h, self.h, self.max_tris), lib().hiprt_last_error().decode()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hiprt_bvh_free(self.h)
            self.h = None


def camera_rays(cam):
    """Host camera rays (8, w*h) float32 for cam = pos, dir, up, fovy, w, h."""
    cam = np.ascontiguousarray(cam[:12], dtype=np.float32)
    w, h = int(cam[10]), int(cam[11])
    out = np.empty((8, w * h), np.float32)
    planes = (c_void_p * 8)(*[out[k].ctypes.data for k in range(8)])
    check(lib().hiprt_camera_rays_host(cam.ctypes.data_as(POINTER(c_float)), planes))
    return out


class Context:
    def __init__(self, device=0):
        self.h = c_void_p()
        check(lib().hiprt_create(device, ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().hiprt_destroy(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()


class DeviceScene:
    def __init__(self, ctx, scene, bvh, bvh_format=BVH2, tri_format=TRI_EXACT, handle=None):
        self.ctx = ctx
        if handle is not None:
            self.h = handle
        else:
            self.h = c_void_p()
            check(lib().hiprt_scene_commit(ctx.h, scene.h, bvh.h, bvh_format, tri_format, ctypes.byref(self.h)))
        nn, ns, root, depth, bf, tf = c_uint32(), c_uint32(), c_int32(), c_uint32(), c_int32(), c_int32()
        check(lib().hiprt_scene_layout(self.h, ctypes.byref(nn), ctypes.byref(ns), ctypes.byref(root),
                                       ctypes.byref(depth), ctypes.byref(bf), ctypes.byref(tf)))
        self.n_nodes, self.n_slots, self.root_ref, self.depth = nn.value, ns.value, root.value, depth.value
        need = c_uint32()
        check(lib().hiprt_scene_stack_need(self.h, ctypes.byref(need)))
        self.stack_need = need.value
        self.bvh_format, self.tri_format = bf.value, tf.value

    def download(self):
        node_bytes = 64 if self.bvh_format == BVH2 else 128
        nodes = np.zeros(self.n_nodes * node_bytes // 4, np.uint32)
        tris = np.zeros(self.n_slots * 12, np.float32)
        ids = np.zeros(self.n_slots, np.uint32)
        check(lib().hiprt_scene_download(self.h, _ptr(nodes), _ptr(tris), _ptr(ids)))
        return nodes, tris, ids

    def refit(self, positions, stream=None):
        """Refit to `positions`, a contiguous (n_vert, 3) float32 CUDA tensor:
        same triangles and topology, new records and boxes."""
        assert positions.dtype.is_floating_point and positions.element_size() == 4 and positions.is_contiguous()
        nv, _ = scene_counts(self)   # 0 for scenes uploaded without a host scene: the library refuses those
        if (nv and tuple(positions.shape) != (nv, 3)) or not positions.is_cuda:
            raise HiprtError(f"refit needs a ({nv}, 3) CUDA tensor, got {tuple(positions.shape)}")
        check(lib().hiprt_scene_refit(self.h, c_void_p(positions.data_ptr()), _stream(stream)))

    def close(self):
        if getattr(self, "h", None):
            lib().hiprt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()


def build_device_scene(ctx, positions, indices, materials=None, material_table=None, max_leaf=4,
                       tri_format=TRI_EXACT, stream=None):
    """Linear BVH built on the device from CUDA tensors: positions (n_vert, 3)
    float32, indices (n_tris, 3) int32, optional per-triangle material ids
    (n_tris,) int32 and a host material table (k, 8) float32.  Returns the
    DeviceScene and the build kernels' time in ms."""
    assert positions.is_cuda and indices.is_cuda and positions.is_contiguous() and indices.is_contiguous()
    assert positions.element_size() == 4 and indices.element_size() == 4
    nv = positions.shape[0] if positions.numel() else 0
    nt = indices.shape[0] if indices.numel() else 0
    table = None if material_table is None else np.ascontiguousarray(material_table, dtype=np.float32).reshape(-1, 8)
    h = c_void_p()
    ms = c_float(0.0)
    check(lib().hiprt_scene_build(ctx.h, c_void_p(positions.data_ptr()) if nv else None, nv,
                                  c_void_p(indices.data_ptr()) if nt else None, nt,
                                  c_void_p(materials.data_ptr()) if materials is not None else None,
                                  table.ctypes.data if table is not None else None,
                                  0 if table is None else table.shape[0], max_leaf, tri_format, ctypes.byref(h),
                                  ctypes.byref(ms), _stream(stream)))
    return DeviceScene(ctx, None, None, handle=h), ms.value


def _rays_struct(rays):
    """rays: torch float32 CUDA tensor of shape (8, n), contiguous."""
    assert rays.dim() == 2 and rays.shape[0] == 8 and rays.is_contiguous()
    return _Rays(*[rays[k].data_ptr() for k in range(8)])


def _hits_struct(hits):
    """hits: dict with t, u, v (float32) and id (int32) CUDA tensors."""
    return _Hits(hits["t"].data_ptr(), hits["u"].data_ptr(), hits["v"].data_ptr(), hits["id"].data_ptr())


def tracer_id(name):
    i = lib().hiprt_tracer_id(name.encode())
    if i < 0:
        raise HiprtError(f"No tracer {name}")
    return i


def _opts(refill_threshold=0, blocks=0, fetch_stats=None, stack_config=0, work_chunks=0, spill_limit=0,
          park_leaves=0):
    return _TracerOpts(refill_threshold, blocks, fetch_stats.data_ptr() if fetch_stats is not None else None,
                       stack_config, work_chunks, spill_limit, park_leaves)


def trace_closest(dscene, rays, hits, tracer="while-while", stream=None, refill_threshold=0, blocks=0,
                  fetch_stats=None, stack_config=0, work_chunks=0, spill_limit=0, park_leaves=0):
    """Closest hits of `rays` (8, n) into `hits`.  fetch_stats: optional int64
    CUDA tensor of 2 that receives inner-node fetches and triangle tests."""
    n = rays.shape[1]
    r, h = _rays_struct(rays), _hits_struct(hits)
    o = _opts(refill_threshold, blocks, fetch_stats, stack_config, work_chunks, spill_limit, park_leaves)
    s = c_void_p(stream) if stream else None
    check(lib().hiprt_trace_closest(dscene.h, ctypes.byref(r), n, ctypes.byref(h), tracer_id(tracer),
                                    ctypes.byref(o), s))


def trace_closest_noalloc(dscene, rays, hits, tracer="persistent-while-while", stream=None, refill_threshold=0,
                          blocks=0, stack_config=0, work_chunks=0):
    """hiprt_trace_closest_noalloc: closest hits with the library's static
    scratch instead of any device allocation."""
    r, h = _rays_struct(rays), _hits_struct(hits)
    o = _opts(refill_threshold, blocks, None, stack_config, work_chunks)
    s = c_void_p(stream) if stream else None
    check(lib().hiprt_trace_closest_noalloc(dscene.h, ctypes.byref(r), rays.shape[1], ctypes.byref(h),
                                            tracer_id(tracer), ctypes.byref(o), s))


def trace_any(dscene, rays, occluded, hits=None, tracer="while-while", stream=None, refill_threshold=0, blocks=0,
              stack_config=0, work_chunks=0, spill_limit=0):
    n = rays.shape[1]
    r, o = _rays_struct(rays), _opts(refill_threshold, blocks, None, stack_config, work_chunks, spill_limit)
    hp = ctypes.byref(_hits_struct(hits)) if hits is not None else None
    occ = c_void_p(occluded.data_ptr()) if occluded is not None else None
    s = c_void_p(stream) if stream else None
    check(lib().hiprt_trace_any(dscene.h, ctypes.byref(r), n, occ, hp, tracer_id(tracer), ctypes.byref(o), s))


def new_hits(n, device="cuda"):
    import torch
    return {"t": torch.empty(n, dtype=torch.float32, device=device),
            "u": torch.empty(n, dtype=torch.float32, device=device),
            "v": torch.empty(n, dtype=torch.float32, device=device),
            "id": torch.empty(n, dtype=torch.int32, device=device)}


def _stream(stream):
    return c_void_p(stream) if stream else None


def camera_rays_device(cam, out, stream=None):
    """GPU camera rays into `out`, a (8, w*h) float32 CUDA tensor."""
    c = np.ascontiguousarray(cam[:12], dtype=np.float32)
    r = _rays_struct(out)
    check(lib().hiprt_camera_rays(c.ctypes.data_as(POINTER(c_float)), ctypes.byref(r), _stream(stream)))


def incoherence_spheres(pos):
    """The two spheres of the camera step's "incoherence" rays for a scene
    with vertex positions `pos` (rt/cuda/rni.cpp:36-72): along the largest
    extent, centred a quarter in from each end, the other axes centred; r_max
    is a quarter of that extent.  float32 arithmetic."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    lo, hi = pos.min(0), pos.max(0)
    size = np.abs(hi - lo)
    f = np.float32
    a = 0 if (size[0] >= size[1] and size[0] >= size[2]) else (1 if size[1] >= size[2] else 2)
    s1 = lo + size / f(2)
    s2 = lo + size / f(2)
    s1[a] = lo[a] + size[a] / f(4)
    s2[a] = hi[a] - size[a] / f(4)
    return s1.astype(np.float32), s2.astype(np.float32), f(size[a] / f(4))


def incoherent_rays_device(sphere1, sphere2, r1, r2, r_max, seed, out, stream=None):
    """hiprt_incoherent_rays into `out`, a (8, n) float32 CUDA tensor."""
    s1 = np.ascontiguousarray(sphere1, dtype=np.float32)
    s2 = np.ascontiguousarray(sphere2, dtype=np.float32)
    r = _rays_struct(out)
    check(lib().hiprt_incoherent_rays(s1.ctypes.data_as(POINTER(c_float)), s2.ctypes.data_as(POINTER(c_float)), r1, r2,
                                      r_max, seed, ctypes.byref(r), out.shape[1], _stream(stream)))


def sort_keys(keys, bits, stream=None):
    """hiprt_sort_keys: the stable order of the uint32 keys (a 1-D CUDA
    tensor, int32 or uint32 storage; left unchanged) by their low `bits`
    bits, as an int32 CUDA tensor of indices."""
    import torch
    n = keys.numel()
    work = keys.contiguous().clone().view(torch.int32)
    perm = torch.empty(n, dtype=torch.int32, device=keys.device)
    ws = torch.empty(max(int(lib().hiprt_sort_keys_workspace(n)), 1), dtype=torch.uint8, device=keys.device)
    check(lib().hiprt_sort_keys(c_void_p(work.data_ptr()), c_void_p(perm.data_ptr()), n, bits,
                                c_void_p(ws.data_ptr()), _stream(stream)))
    return perm


def bounce_cosine(dscene, rays, hits, out, seed, level, stream=None):
    """Cosine-weighted bounce rays off `hits` of `rays` into `out` (both (8, n))."""
    n = rays.shape[1]
    r, h, o = _rays_struct(rays), _hits_struct(hits), _rays_struct(out)
    check(lib().hiprt_bounce_cosine(dscene.h, ctypes.byref(r), ctypes.byref(h), n, seed, level, ctypes.byref(o),
                                    _stream(stream)))


def shadow_rays(dscene, rays, hits, out, light, stream=None):
    """Shadow rays from `hits` of `rays` towards the point `light` into `out`
    ((8, n) each); misses give the disabled ray."""
    n = rays.shape[1]
    r, h, o = _rays_struct(rays), _hits_struct(hits), _rays_struct(out)
    lp = np.ascontiguousarray(light[:3], dtype=np.float32)
    check(lib().hiprt_shadow_rays(dscene.h, ctypes.byref(r), ctypes.byref(h), n,
                                  lp.ctypes.data_as(POINTER(c_float)), ctypes.byref(o), _stream(stream)))


DIRECT_MODES = {"uniform": 0, "cosine": 1, "light": 2}


def set_lights(dscene, points=None):
    """Light list of the direct-lighting steps: the scene's emissive triangles,
    then point lights (k, 6) pos + col.  Returns the number of lights."""
    pts = None if points is None else np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 6)
    n = c_uint32()
    check(lib().hiprt_scene_set_lights(dscene.h, _ptr(pts) if pts is not None else None,
                                       0 if pts is None else len(pts), ctypes.byref(n)))
    return n.value


def sample_direct(dscene, mode, cam_rays, cam_hits, seed, out, pdf, fb, light_col=None, stream=None):
    """hiprt_sample_direct: mode "uniform", "cosine" or "light"; out (8, n),
    pdf (n,), fb (n, 4), light_col (n, 3) in light mode."""
    check(lib().hiprt_sample_direct(dscene.h, DIRECT_MODES[mode], ctypes.byref(_rays_struct(cam_rays)),
                                    ctypes.byref(_hits_struct(cam_hits)), cam_rays.shape[1], seed,
                                    ctypes.byref(_rays_struct(out)), c_void_p(pdf.data_ptr()),
                                    _dp(light_col), c_void_p(fb.data_ptr()), _stream(stream)))


def integrate_light(dscene, cam_rays, cam_hits, shadow_rays, occluded, light, col, fb, stream=None):
    """hiprt_integrate_light: one point light at `light` of colour `col`."""
    lp = np.ascontiguousarray(light[:3], dtype=np.float32)
    c = np.ascontiguousarray(col[:3], dtype=np.float32)
    check(lib().hiprt_integrate_light(dscene.h, ctypes.byref(_rays_struct(cam_rays)),
                                      ctypes.byref(_hits_struct(cam_hits)), ctypes.byref(_rays_struct(shadow_rays)),
                                      c_void_p(occluded.data_ptr()), cam_rays.shape[1],
                                      lp.ctypes.data_as(POINTER(c_float)), c.ctypes.data_as(POINTER(c_float)),
                                      c_void_p(fb.data_ptr()), _stream(stream)))


def integrate_dir_sample(dscene, cam_rays, cam_hits, shadow_rays, shadow_hits, pdf, fb, stream=None):
    check(lib().hiprt_integrate_dir_sample(dscene.h, ctypes.byref(_rays_struct(cam_rays)),
                                           ctypes.byref(_hits_struct(cam_hits)),
                                           ctypes.byref(_rays_struct(shadow_rays)),
                                           ctypes.byref(_hits_struct(shadow_hits)), c_void_p(pdf.data_ptr()),
                                           cam_rays.shape[1], c_void_p(fb.data_ptr()), _stream(stream)))


def integrate_light_sample(dscene, cam_rays, cam_hits, shadow_rays, occluded, pdf, light_col, fb, stream=None):
    check(lib().hiprt_integrate_light_sample(dscene.h, ctypes.byref(_rays_struct(cam_rays)),
                                             ctypes.byref(_hits_struct(cam_hits)),
                                             ctypes.byref(_rays_struct(shadow_rays)),
                                             c_void_p(occluded.data_ptr()), c_void_p(pdf.data_ptr()),
                                             c_void_p(light_col.data_ptr()), cam_rays.shape[1],
                                             c_void_p(fb.data_ptr()), _stream(stream)))


class PathTracer:
    """Wavefront diffuse path tracer over a device scene (plain or batched)."""

    def __init__(self, dscene, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True, tracer=None,
                 stack_config=0, work_chunks=0, refill_threshold=0, blocks=0, sort_rays=True, samples=1,
                 emitters_reflect=False, jitter=True, park_leaves=0, regen_pool=0, refill_coherent=0):
        """regen_pool k > 0: path regeneration over k path slots per pixel
        (hiprt_pt_params.regen_pool) instead of all `samples` in flight.
        refill_coherent: the dynamic tracer's refill threshold on the camera
        level and the first bounce with samples in flight (0 = 16)."""
        self.dscene = dscene
        c = np.ascontiguousarray(cam[:12], dtype=np.float32)
        self.w, self.h = int(c[10]), int(c[11])
        self.levels = max_depth + 1
        tid = -1 if tracer is None else tracer_id(tracer)
        p = _PtParams(max_depth, rr_start, rr_survive, seed, tid, 1 if batched else 0, stack_config, work_chunks,
                      refill_threshold, blocks, 0 if sort_rays else 1, samples, 1 if emitters_reflect else 0,
                      park_leaves, 0 if jitter else 1, regen_pool, refill_coherent)
        self.h_ = c_void_p()
        check(lib().hiprt_pt_create(dscene.h, c.ctypes.data_as(POINTER(c_float)), ctypes.byref(p),
                                    ctypes.byref(self.h_)))

    def render(self, frame=0, stream=None):
        check(lib().hiprt_pt_render(self.h_, frame, _stream(stream)))

    def ray_counts(self):
        out = (c_uint32 * self.levels)()
        check(lib().hiprt_pt_ray_counts(self.h_, out, self.levels))
        return np.array(out, dtype=np.uint64)

    def ray_totals(self):
        out = (c_uint64 * self.levels)()
        check(lib().hiprt_pt_ray_totals(self.h_, out, self.levels))
        return np.array(out, dtype=np.uint64)

    def profile(self, frame=0, stream=None):
        """Per-level trace milliseconds, (inner-node fetches, triangle tests), live rays."""
        ms = (c_float * self.levels)()
        st = (c_uint64 * (2 * self.levels))()
        rays = (c_uint32 * self.levels)()
        check(lib().hiprt_pt_profile(self.h_, frame, _stream(stream), ms, st, rays, self.levels))
        return (np.array(ms, np.float64), np.array(st, np.uint64).reshape(-1, 2), np.array(rays, np.uint64))

    def memory(self):
        """Device bytes the path tracer holds (hiprt_pt_memory)."""
        b = c_uint64()
        check(lib().hiprt_pt_memory(self.h_, ctypes.byref(b)))
        return int(b.value)

    def image(self):
        img = np.empty((self.h, self.w, 3), np.float32)
        check(lib().hiprt_pt_image(self.h_, c_void_p(img.ctypes.data)))
        return img

    def reset(self):
        check(lib().hiprt_pt_reset(self.h_, None))

    def close(self):
        if getattr(self, "h_", None):
            lib().hiprt_pt_destroy(self.h_)
            self.h_ = None

    def __del__(self):
        if _lib is not None:
            self.close()


def pt_autotune(dscene, dscene4, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, max_samples=64,
                trial_frames=2, sort_rays=True):
    """hiprt_pt_autotune: the batched path tracer's configuration for this
    scene and camera (dscene4: the 4-wide layout or None).  Returns the
    PathTracer keyword arguments of the winner and the measurement."""
    c = np.ascontiguousarray(cam[:12], dtype=np.float32)
    base = _PtParams(max_depth, rr_start, rr_survive, seed, -1, 1, 0, 0, 0, 0, 0 if sort_rays else 1, 1, 0, 0, 0)
    res = _PtTuneResult()
    check(lib().hiprt_pt_autotune(dscene.h, dscene4.h if dscene4 is not None else None,
                                  c.ctypes.data_as(POINTER(c_float)), ctypes.byref(base), max_samples, trial_frames,
                                  ctypes.byref(res)))
    kw = {"tracer": TRACERS[res.tracer], "stack_config": res.stack_config, "samples": res.samples,
          "park_leaves": res.park_leaves}
    return kw, {"layout": "bvh4" if res.bvh_format == BVH4 else "bvh2", "mrays_per_s": round(res.mrays_per_s, 1),
                "candidates": res.candidates}


def _dp(t):
    return c_void_p(t.data_ptr()) if t is not None else None


def scene_counts(dscene):
    nv, nt = c_uint32(), c_uint32()
    check(lib().hiprt_scene_counts(dscene.h, ctypes.byref(nv), ctypes.byref(nt)))
    return nv.value, nt.value


def adjoint_closest(dscene, rays, hits, grad_t, grad_u, grad_v, grad_vertices=None, grad_origin=None,
                    grad_dir=None, grad_corners=None, stream=None):
    """Backward of the closest hit with the hit set of `hits` frozen.  Upstream
    gradients are float32 (n,) CUDA tensors; outputs (all optional): vertex
    gradients (n_vert, 3) accumulated, origin / direction gradients (n, 3) and
    per-ray corner gradients (n, 9), overwritten."""
    n = rays.shape[1]
    r, h = _rays_struct(rays), _hits_struct(hits)
    check(lib().hiprt_adjoint_closest(dscene.h, ctypes.byref(r), ctypes.byref(h), n, _dp(grad_t), _dp(grad_u),
                                      _dp(grad_v), _dp(grad_vertices), _dp(grad_origin), _dp(grad_dir),
                                      _dp(grad_corners), _stream(stream)))


def trace_batch(entries, any=False, tracer="dynamic-while-while", stream=None, refill_threshold=0, blocks=0,
                stack_config=0, work_chunks=0):
    """entries: list of (dscene, rays (8,n), hits dict or None, occluded tensor or None)."""
    arr = (_BatchEntry * len(entries))()
    for k, (ds, rays, hits, occ) in enumerate(entries):
        arr[k].scene = ds.h
        arr[k].rays = _rays_struct(rays)
        arr[k].hits = _hits_struct(hits) if hits is not None else _Hits(None, None, None, None)
        arr[k].occluded = occ.data_ptr() if occ is not None else None
        arr[k].n = rays.shape[1]
    o = _opts(refill_threshold, blocks, None, stack_config, work_chunks)
    check(lib().hiprt_trace_batch(arr, len(entries), 1 if any else 0, tracer_id(tracer), ctypes.byref(o),
                                  _stream(stream)))


def autotune(ctx, scene, bvh, sample_rays, any=False):
    """Returns (result dict, DeviceScene committed in the chosen layout)."""
    res = _TuneResult()
    h = c_void_p()
    r = _rays_struct(sample_rays)
    check(lib().hiprt_autotune(ctx.h, scene.h, bvh.h, ctypes.byref(r), sample_rays.shape[1], 1 if any else 0,
                               ctypes.byref(res), ctypes.byref(h)))
    ds = DeviceScene(ctx, None, None, handle=h)
    out = {"tracer": TRACERS[res.tracer], "refill_threshold": res.refill_threshold, "blocks": res.blocks,
           "bvh_format": "bvh4" if res.bvh_format == BVH4 else "bvh2", "mrays_per_s": res.mrays_per_s,
           "candidates": res.candidates, "stack_config": res.stack_config, "work_chunks": res.work_chunks,
           "park_leaves": res.park_leaves}
    return out, ds
