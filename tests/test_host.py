"""Host side of the product: generators, the restated BVH builders, the
flattener and the C ABI's host entry points.  No GPU is touched."""
import ctypes
import os
import re

import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, uniform_sphere_rays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- C ABI surface

def test_library_exports_every_declared_symbol():
    L = hiprt.lib()
    names = hiprt.exported_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_errors_are_reported_not_thrown():
    L = hiprt.lib()
    h = ctypes.c_void_p()
    assert L.hiprt_scene_generate(9, ctypes.byref(h), None) == 1
    assert b"config" in L.hiprt_last_error()
    assert L.hiprt_scene_load_obj(b"/nonexistent.obj", ctypes.byref(h)) == 4
    # the reference links "default" and "find closest hits" to while-while
    # (rt/cuda/platform.cpp:55-56)
    assert hiprt.tracer_id("default") == hiprt.tracer_id("while-while") == 0
    assert hiprt.tracer_id("find closest hits") == 0
    assert hiprt.tracer_id("default-alpha") == hiprt.tracer_id("while-while-alpha")
    with pytest.raises(hiprt.HiprtError):
        hiprt.tracer_id("no-such-tracer")
    # the reference's other loop structures name the kernel of the same traversal
    for alias, canonical in (("simple", "while-while"), ("if-if", "while-while"),
                             ("persistent-if-if", "persistent-while-while"),
                             ("speculative-while-while", "persistent-speculative-while-while")):
        assert hiprt.tracer_id(alias) == hiprt.tracer_id(canonical)
        assert hiprt.tracer_id(alias + "-alpha") == hiprt.tracer_id(canonical + "-alpha")
        assert L.hiprt_tracer_name(hiprt.tracer_id(alias)).decode() == canonical
    for bad_name in ("if-if-", "if-ifalpha", "-alpha", "simple-alpha-alpha", "if"):
        with pytest.raises(hiprt.HiprtError):
            hiprt.tracer_id(bad_name)
    bad = np.array([[0, 1, 5]], np.uint32)
    with pytest.raises(hiprt.HiprtError, match="out of range"):
        hiprt.HostScene.from_arrays(np.zeros((3, 3), np.float32), bad)


# ---------------------------------------------------------------- generators

def test_workload_sizes():
    assert hiprt.HostScene.generate(1).n_tris == 8192
    assert hiprt.HostScene.generate(2).n_tris == 74068
    assert hiprt.HostScene.generate(3).n_tris == 262144


def test_random_scene_statistics_and_determinism():
    a = hiprt.HostScene.random(8192, 0.02, 1)
    b = hiprt.HostScene.random(8192, 0.02, 1)
    pa, ia, _ = a.arrays()
    pb, _, _ = b.arrays()
    assert np.array_equal(pa, pb)
    tri = pa[ia]                                   # (n, 3, 3)
    cent = tri.mean(axis=1)
    off = (tri - cent[:, None, :]).reshape(-1, 3)
    # centroids U[0,1]^3 (within the per-vertex noise), offsets ~ N(0, 0.02^2)
    assert abs(cent.mean() - 0.5) < 0.01 and cent.min() > -0.1 and cent.max() < 1.1
    # offsets from the sample centroid have variance 2/3 sigma^2
    assert abs(off.std() - 0.02 * np.sqrt(2 / 3)) < 0.001


def test_obj_round_trip(tmp_path):
    sc = hiprt.HostScene.generate(1)
    p = tmp_path / "c1.obj"
    sc.write_obj(p)
    back = hiprt.HostScene.load_obj(p)
    a, b = sc.arrays(), back.arrays()
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))   # %.9g round-trips floats
    assert np.array_equal(a[1], b[1])


def test_obj_reader_fans_polygons(tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1/1 2/2 3/3 4/4\n")
    pos, idx, _ = hiprt.HostScene.load_obj(p).arrays()
    assert idx.tolist() == [[0, 1, 2], [0, 2, 3]]


def test_obj_materials_from_a_library(tmp_path):
    """mtllib materials resolved as RTGI's loader does: albedo = Kd^2.2, or
    Ks^2.2 when Kd's luma is ~0; emission = Ke.  As assimp's OBJ importer
    lists them: "DefaultMaterial" first, the library's in definition order,
    then names "usemtl" uses without a definition (Kd 0.6 grey, as assimp's
    material defaults)."""
    (tmp_path / "lib.mtl").write_text("newmtl red\nKd 0.8 0.1 0.1\n\nnewmtl lamp\nKd 0 0 0\nKs 0.5 0.25 0.5\n"
                                      "Ke 4 3 2\nnewmtl plain\n")
    (tmp_path / "m.obj").write_text(
        "mtllib lib.mtl\n" + "".join(f"v {i} {i % 2} {i % 3}\n" for i in range(7)) +
        "usemtl red\nf 1 2 3\nusemtl lamp\nf 4 5 6\nusemtl nothing\nf 1 3 5\nf 2 4 6 7\n")
    sc = hiprt.HostScene.load_obj(tmp_path / "m.obj")
    _, idx, mat = sc.arrays()
    table = sc.materials()
    assert sc.material_names() == ["DefaultMaterial", "red", "lamp", "plain", "nothing"]
    assert mat.tolist() == [1, 2, 4, 4, 4]
    assert table.shape == (5, 8)
    f32 = np.float32
    want_red = np.array([0.8, 0.1, 0.1], f32) ** f32(2.2)
    want_lamp = np.array([0.5, 0.25, 0.5], f32) ** f32(2.2)
    grey = f32(0.6) ** f32(2.2)
    assert np.allclose(table[1, :3], want_red, rtol=1e-6) and (table[1, 4:7] == 0).all()
    assert np.allclose(table[2, :3], want_lamp, rtol=1e-6) and table[2, 4:7].tolist() == [4, 3, 2]
    for k in (0, 3, 4):   # the default, a definition without Kd, an undefined name
        assert np.allclose(table[k, :3], grey, rtol=1e-6) and (table[k, 4:7] == 0).all()
    # without any library: one default material
    (tmp_path / "p.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    sc = hiprt.HostScene.load_obj(tmp_path / "p.obj")
    assert sc.material_names() == ["DefaultMaterial"] and sc.arrays()[2].tolist() == [0]
    assert np.allclose(sc.materials()[0, :3], grey, rtol=1e-6)


def test_material_names_and_edits(tmp_path):
    """Material names as "material select" finds them: "DefaultMaterial",
    then the library's in definition order, then undefined "usemtl" names;
    albedo and emission can be set per material."""
    (tmp_path / "lib.mtl").write_text("newmtl red\nKd 0.8 0.1 0.1\nnewmtl lamp\nKd 0.5 0.5 0.5\n")
    (tmp_path / "m.obj").write_text("mtllib lib.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl lamp\nf 1 2 3\n")
    sc = hiprt.HostScene.load_obj(tmp_path / "m.obj")
    assert sc.material_names() == ["DefaultMaterial", "red", "lamp"]
    sc.set_material(2, emissive=[5, 4, 3])
    sc.set_material(1, albedo=[0.25, 0.5, 0.75])
    t = sc.materials()
    assert t[2, 4:7].tolist() == [5, 4, 3] and t[1, :3].tolist() == [0.25, 0.5, 0.75]
    assert np.allclose(t[2, :3], 0.5 ** 2.2)
    with pytest.raises(hiprt.HiprtError, match="material id"):
        sc.set_material(3, albedo=[1, 1, 1])
    (tmp_path / "l.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl m1\nf 1 2 3\n")
    assert hiprt.HostScene.load_obj(tmp_path / "l.obj").material_names() == ["DefaultMaterial", "m1"]
    assert hiprt.HostScene.generate(1).material_names() == [""]


def test_obj_missing_library_falls_back_to_the_obj_name(tmp_path):
    """A "mtllib" that cannot be opened falls back to the OBJ file's name with
    "mtl" for "obj" (assimp), and loads without a library when that is
    missing too."""
    body = "mtllib gone.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl warm\nf 1 2 3\n"
    (tmp_path / "x.obj").write_text(body)
    (tmp_path / "x.mtl").write_text("newmtl warm\nKd 1 0.5 0.25\n")
    sc = hiprt.HostScene.load_obj(tmp_path / "x.obj")
    assert sc.material_names() == ["DefaultMaterial", "warm"]
    assert np.allclose(sc.materials()[1, :3], np.array([1, 0.5, 0.25], np.float32) ** np.float32(2.2))
    (tmp_path / "y.obj").write_text(body)
    sc = hiprt.HostScene.load_obj(tmp_path / "y.obj")
    assert sc.material_names() == ["DefaultMaterial", "warm"]   # undefined: default values
    assert np.allclose(sc.materials()[1, :3], np.float32(0.6) ** np.float32(2.2))


def _tri_lines(n):
    """Vertices of n separate triangles, triangle k at x = k."""
    return "".join(f"v {k} 0 0\nv {k + 0.5} 0 0\nv {k} 0.5 0\n" for k in range(n))


def _tri_order(sc):
    pos, idx, mat = sc.arrays()
    return [int(round(float(pos[i[0], 0]))) for i in idx], mat.tolist()


def test_obj_triangle_order_follows_assimp_objects_and_meshes(tmp_path):
    """Triangle order and materials as the reference sees them through
    assimp's OBJ importer: objects in creation order, each object's meshes in
    creation order.  "o" re-opens an object of the same name; "usemtl" with
    another material starts a mesh in the current object once the current
    mesh has faces; faces before an object's first "usemtl" take that
    material; "g" opens a new object whenever the group name changes."""
    (tmp_path / "lib.mtl").write_text("newmtl a\nKd 1 0 0\nnewmtl b\nKd 0 1 0\n")

    def f(k):
        return f"f {3 * k + 1} {3 * k + 2} {3 * k + 3}\n"

    # o A (a) t0 | o B (b) t1 | o A, usemtl a: a new mesh in A -> t2, t3
    (tmp_path / "o.obj").write_text("mtllib lib.mtl\n" + _tri_lines(4) + "o A\nusemtl a\n" + f(0) +
                                    "o B\nusemtl b\n" + f(1) + "o A\nusemtl a\n" + f(2) + f(3))
    assert _tri_order(hiprt.HostScene.load_obj(tmp_path / "o.obj")) == ([0, 2, 3, 1], [1, 1, 1, 2])
    # re-opening without a material switch: faces stay in the current mesh (B's)
    (tmp_path / "r.obj").write_text("mtllib lib.mtl\n" + _tri_lines(3) + "o A\nusemtl a\n" + f(0) +
                                    "o B\nusemtl b\n" + f(1) + "o A\n" + f(2))
    assert _tri_order(hiprt.HostScene.load_obj(tmp_path / "r.obj")) == ([0, 1, 2], [1, 2, 2])
    # a face before the first usemtl of its mesh takes that material
    (tmp_path / "u.obj").write_text("mtllib lib.mtl\n" + _tri_lines(3) + "o A\n" + f(0) + "usemtl b\n" + f(1) +
                                    "usemtl a\n" + f(2))
    assert _tri_order(hiprt.HostScene.load_obj(tmp_path / "u.obj")) == ([0, 1, 2], [2, 2, 1])
    # groups: a new object per group change, even for a name seen before
    (tmp_path / "g.obj").write_text("mtllib lib.mtl\n" + _tri_lines(3) + "g A\nusemtl a\n" + f(0) +
                                    "g B\nusemtl b\n" + f(1) + "g A\nusemtl a\n" + f(2))
    assert _tri_order(hiprt.HostScene.load_obj(tmp_path / "g.obj")) == ([0, 1, 2], [1, 2, 1])
    # usemtl before any object; faces with no object open "defaultobject"
    (tmp_path / "d.obj").write_text("mtllib lib.mtl\n" + _tri_lines(2) + "usemtl b\n" + f(0) + f(1))
    assert _tri_order(hiprt.HostScene.load_obj(tmp_path / "d.obj")) == ([0, 1], [2, 2])


def test_obj_concave_quad_splits_from_its_concave_corner(tmp_path):
    """A quad is split along the diagonal through its concave corner (assimp's
    quad rule): the two triangles then tile the quad, with the quad's
    orientation and its area.  A fan from corner 0 would not: one of its
    triangles would lie outside the dart."""
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 2 1 0\nv 0 2 0\nv 0.5 1 0\nf 1 2 3 4\n"
                                    "v 10 0 0\nv 12 0 0\nv 12 2 0\nv 10 2 0\nf 5 6 7 8\n")
    pos, idx, _ = hiprt.HostScene.load_obj(tmp_path / "q.obj").arrays()
    p = pos.astype(np.float64)
    signed = [0.5 * np.cross(p[b] - p[a], p[c] - p[a])[2] for a, b, c in idx]
    assert len(idx) == 4 and all(s > 0 for s in signed)
    assert abs(signed[0] + signed[1] - 1.5) < 1e-6   # the dart's area (shoelace)
    assert abs(signed[2] + signed[3] - 4.0) < 1e-6
    assert idx[0].tolist() == [3, 0, 1] and idx[1].tolist() == [3, 1, 2]
    assert idx[2].tolist() == [4, 5, 6] and idx[3].tolist() == [4, 6, 7]   # convex: from corner 0


def test_add_obj_appends_under_a_model_matrix(tmp_path):
    """hiprt_scene_add_obj (scene::add): models accumulate; material ids of a
    later model follow the scene's materials and its names carry the
    "<name>/" prefix; positions move by the model matrix and normals by its
    inverse transpose (normalized), before the winding check -- so under a
    mirroring matrix b and c swap and geometric and shading normals still
    agree."""
    (tmp_path / "lib.mtl").write_text("newmtl red\nKd 1 0 0\n")
    (tmp_path / "n.obj").write_text("mtllib lib.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0.6 0 0.8\nusemtl red\n"
                                    "f 1//1 2//1 3//1\n")
    (tmp_path / "p.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    sc = hiprt.HostScene.add_obj(tmp_path / "p.obj")
    assert sc.normals() is None
    M = np.array([[-2, 0, 0, 5], [0, 1, 0, 1], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)   # mirror x, scale 2, move
    hiprt.HostScene.add_obj(tmp_path / "n.obj", name="lamp", trafo=M, scene=sc)
    pos, idx, mat = sc.arrays()
    assert sc.material_names() == ["DefaultMaterial", "lamp/DefaultMaterial", "lamp/red"]
    assert mat.tolist() == [0, 2]
    nrm = sc.normals()
    assert nrm is not None and len(nrm) == len(pos)
    # the first model gained flat normals: its face normal, +z
    assert np.allclose(nrm[idx[0]], [[0, 0, 1]] * 3)
    # the second: corners moved, b and c swapped by the mirror
    want = (M[:3, :3] @ np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32).T).T + M[:3, 3]
    assert np.allclose(pos[idx[1]], want)
    n = np.linalg.inv(M[:3, :3].astype(np.float64)).T @ np.array([0.6, 0, 0.8])
    assert np.allclose(nrm[idx[1]], [n / np.linalg.norm(n)] * 3, atol=1e-6)
    a, b, c = pos[idx[1]].astype(np.float64)
    assert np.dot(np.cross(b - a, c - a), nrm[idx[1][0]]) > 0
    with pytest.raises(hiprt.HiprtError, match="cannot open"):
        hiprt.HostScene.add_obj(tmp_path / "none.obj", scene=sc)
    assert sc.n_tris == 2


def test_obj_winding_follows_shading_normals(tmp_path):
    """A triangle whose geometric normal points away from the mean of its
    corner normals has b and c swapped (libgi/scene.cpp:271-277); one whose
    normals agree, or that has no normals, keeps its order.  Every (v, vn)
    corner is a vertex of its own carrying its normal; the corners of the face
    without normals get the face's normal."""
    (tmp_path / "n.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0 -1\n"
        "f 1//1 2//1 3//1\n"            # geometric normal +z, shading +z: kept
        "f 1//2 2//2 3//2\n"            # shading -z: swapped
        "f -3//-1 -2//-1 -1//-1\n"      # relative indices, shading -z: swapped
        "f 1 2 3\n")                    # no normals: kept, face normal +z
    sc = hiprt.HostScene.load_obj(tmp_path / "n.obj")
    pos, idx, _ = sc.arrays()
    nrm = sc.normals()
    P = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    kept, swapped = [P[0], P[1], P[2]], [P[0], P[2], P[1]]
    assert pos[idx].tolist() == [kept, swapped, swapped, kept]
    up, down = [0, 0, 1], [0, 0, -1]
    assert nrm[idx].tolist() == [[up] * 3, [down] * 3, [down] * 3, [up] * 3]
    assert len(pos) == 3 + 3 + 3 + 3   # the file's vertices, then one per distinct (v, vn) or face corner
    assert hiprt.HostScene.load_obj(tmp_path / "n.obj").normals() is not None


def test_scene_positions_setter():
    """hiprt_scene_set_positions replaces every vertex position and refuses a
    count that differs from the scene's."""
    sc = hiprt.HostScene.random(10, 0.1, 3)
    pos, idx, _ = sc.arrays()
    moved = pos * np.float32(2) + np.float32(1)
    sc.set_positions(moved)
    p2, i2, _ = sc.arrays()
    assert np.array_equal(p2, moved) and np.array_equal(i2, idx)
    with pytest.raises(hiprt.HiprtError, match="position count"):
        sc.set_positions(moved[:-1])


def test_scene_normals_round_trip_and_counts(tmp_path):
    """hiprt_scene_set_normals takes one normal per vertex (or none); a scene
    without normals reports none, and an OBJ without "vn" has none."""
    sc = hiprt.HostScene.generate(1)
    pos, _, _ = sc.arrays()
    assert sc.normals() is None
    n = np.tile(np.array([[0, 1, 0]], np.float32), (len(pos), 1))
    sc.set_normals(n)
    assert np.array_equal(sc.normals(), n)
    with pytest.raises(hiprt.HiprtError):
        sc.set_normals(n[:-1])
    sc.set_normals(None)
    assert sc.normals() is None
    (tmp_path / "p.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert hiprt.HostScene.load_obj(tmp_path / "p.obj").normals() is None


def test_obj_texture_coordinates_per_corner(tmp_path):
    """"vt" coordinates belong to face corners: each distinct (position, vt)
    pair becomes its own vertex after the file's vertices (positions copied),
    corners without one keep the shared vertex with (0, 0); relative vt
    indices count back from the last "vt"; the winding swap moves the
    coordinates with their corners."""
    (tmp_path / "t.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\n"
        "vt 0.25 0.5\nvt 0.75 0.5\nvt 0.5 1.5\n"
        "f 1/1 2/2 3/3\n"            # three new vertices 4, 5, 6
        "f 2/2 4 3/-1\n")            # (2, vt2) reused, 4 plain, (3, vt3) reused
    sc = hiprt.HostScene.load_obj(tmp_path / "t.obj")
    pos, idx, _ = sc.arrays()
    tc = sc.texcoords()
    assert len(pos) == 7 and idx.tolist() == [[4, 5, 6], [5, 3, 6]]
    assert np.array_equal(pos[4:], pos[[0, 1, 2]])
    assert tc.tolist() == [[0, 0], [0, 0], [0, 0], [0, 0], [0.25, 0.5], [0.75, 0.5], [0.5, 1.5]]
    assert sc.normals() is None
    # with normals in the file: the swap moves coordinates and normals with their corners
    (tmp_path / "w.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25 0.5\nvt 0.75 0.5\nvt 0.5 1.5\nvn 0 0 -1\n"
        "f 1/1/1 2/2/1 3/3/1\n")     # shading -z: b and c swap
    sc = hiprt.HostScene.load_obj(tmp_path / "w.obj")
    pos, idx, _ = sc.arrays()
    assert idx.tolist() == [[3, 4, 5]]
    assert pos[idx[0]].tolist() == [[0, 0, 0], [0, 1, 0], [1, 0, 0]]
    assert sc.texcoords()[idx[0]].tolist() == [[0.25, 0.5], [0.5, 1.5], [0.75, 0.5]]
    assert sc.normals()[idx[0]].tolist() == [[0, 0, -1]] * 3
    (tmp_path / "b.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2/2 3/1\n")
    with pytest.raises(hiprt.HiprtError, match="texture coordinate index"):
        hiprt.HostScene.load_obj(tmp_path / "b.obj")


def test_texture_api_checks_its_arguments():
    sc = hiprt.HostScene.random(10, 0.1, 1)
    with pytest.raises(hiprt.HiprtError, match="vertex count"):
        sc.set_texcoords(np.zeros((7, 2)))
    with pytest.raises(hiprt.HiprtError, match="material id"):
        sc.set_texture(5, np.ones((2, 2, 4)))
    sc.set_texcoords(np.full((30, 2), 0.5))
    assert (sc.texcoords() == 0.5).all()
    sc.set_texture(0, np.ones((2, 3, 4)))
    sc.set_texture(0, None)


def test_obj_rejects_bad_normal_index(tmp_path):
    (tmp_path / "b.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//2\n")
    with pytest.raises(hiprt.HiprtError, match="normal index"):
        hiprt.HostScene.load_obj(tmp_path / "b.obj")


# ---------------------------------------------------------------- BVH builders

@pytest.mark.parametrize("split", [hiprt.SPLIT_OM, hiprt.SPLIT_SM, hiprt.SPLIT_SAH])
@pytest.mark.parametrize("esc", [False, True])
@pytest.mark.parametrize("max_tris", [1, 4, 7])
def test_builder_invariants(split, esc, max_tris):
    sc = hiprt.HostScene.random(3000, 0.03, 5)
    b = hiprt.Bvh(sc, split=split, max_tris=max_tris, planes=16, esc=esc)
    code, msg = b.validate()
    assert code == 0, msg
    # root is node 0 and node ids are a pre-order: children after parents
    inner = b.nodes["link_r"] >= 0
    ids = np.nonzero(inner)[0]
    assert b.root == 0
    assert np.all(b.nodes["link_l"][ids] > ids) and np.all(b.nodes["link_r"][ids] > ids)
    leaves = ~inner
    assert np.all(-b.nodes["link_r"][leaves] <= max_tris)
    if esc:
        assert len(b.index) >= sc.n_tris
    else:
        assert np.array_equal(np.sort(b.index), np.arange(sc.n_tris))


def _box_contains(outer, inner):
    return np.all(outer[0] <= inner[0]) and np.all(outer[1] >= inner[1])


def test_sah_children_enclose_their_triangles():
    # without ESC every inner node's child box holds its subtree's triangle boxes
    sc = hiprt.HostScene.random(2000, 0.03, 2)
    pos, idx, _ = sc.arrays()
    b = hiprt.Bvh(sc, esc=False)
    tri_box = np.stack([pos[idx].min(axis=1), pos[idx].max(axis=1)], axis=1)   # (n, 2, 3)

    def subtree_box(nid):
        nd = b.nodes[nid]
        if nd["link_r"] < 0:
            s = b.index[-nd["link_l"]: -nd["link_l"] - nd["link_r"]]
            return np.stack([tri_box[s, 0].min(0), tri_box[s, 1].max(0)])
        l, r = subtree_box(nd["link_l"]), subtree_box(nd["link_r"])
        assert _box_contains(nd["box_l"], l) and _box_contains(nd["box_r"], r)
        return np.stack([np.minimum(l[0], r[0]), np.maximum(l[1], r[1])])

    whole = subtree_box(b.root)
    assert np.allclose(whole[0], tri_box[:, 0].min(0)) and np.allclose(whole[1], tri_box[:, 1].max(0))


def test_sah_split_separates_clusters():
    # two clusters of 8 tiny triangles far apart on x: the root split must
    # put each cluster in its own child
    rng = np.random.default_rng(0)
    tris = []
    for cx in (0.0, 10.0):
        for _ in range(8):
            c = np.array([cx, 0, 0]) + rng.uniform(-0.5, 0.5, 3)
            tris.append(c + rng.normal(0, 0.05, (3, 3)))
    pos = np.concatenate(tris).astype(np.float32)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    sc = hiprt.HostScene.from_arrays(pos, idx)
    b = hiprt.Bvh(sc, max_tris=4, esc=False)
    root = b.nodes[b.root]
    assert root["box_l"][1][0] < 5 and root["box_r"][0][0] > 5


def test_n_minus_one_split_is_treated_as_degenerate():
    # 5 triangles: 4 clustered at x~0 and one at x~10.  Every SAH plane puts
    # the four on the left and the loner on the right -- mid == end-1 -- which
    # the reference treats as degenerate (bvh-ctor.h:441) and replaces by the
    # object median (2 | 3), although (4 | 1) would be cheaper.
    pts = [[0, 0, 0], [0.1, 0, 0], [0.2, 0, 0], [0.3, 0, 0], [10, 0, 0]]
    tris = [np.array(p) + np.array([[0, 0, 0], [0.05, 0.1, 0], [0, 0.1, 0.05]]) for p in pts]
    pos = np.concatenate(tris).astype(np.float32)
    idx = np.arange(15, dtype=np.uint32).reshape(-1, 3)
    sc = hiprt.HostScene.from_arrays(pos, idx)
    b = hiprt.Bvh(sc, max_tris=1, planes=16, esc=False)
    root = b.nodes[b.root]
    left = b.nodes[root["link_l"]]
    # object median of 5: left child gets 2 primitives (an inner node of 2 leaves)
    assert left["link_r"] >= 0
    assert b.nodes[left["link_l"]]["link_r"] < 0 and b.nodes[left["link_r"]]["link_r"] < 0


def test_root_leaf_for_tiny_scene():
    sc = hiprt.HostScene.random(3, 0.02, 1)
    b = hiprt.Bvh(sc, max_tris=4)
    assert len(b.nodes) == 1 and b.nodes[0]["link_r"] == -3 and b.nodes[0]["link_l"] == 0
    pos, idx, _ = sc.arrays()
    rays = uniform_sphere_rays(3000, 4)
    a = orc.trace_closest(pos, idx, b.nodes, b.root, b.index, rays)
    s = orc.seq_closest(pos, idx, rays)
    assert np.array_equal(a[3], s[3]) and np.array_equal(a[0], s[0])


def test_empty_scene():
    sc = hiprt.HostScene.from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    b = hiprt.Bvh(sc)
    assert len(b.nodes) == 0 and len(b.index) == 0


def test_esc_fragments_cover_triangles():
    # ESC boxes bound fragments, not triangles: every point of a triangle lies
    # in the union of its fragments' leaf boxes.  Checked by sampling points
    # and finding, for each, a leaf that lists the triangle and whose parent
    # box contains the point.
    sc = hiprt.HostScene.generate(2)   # long wall strips get clipped
    pos, idx, _ = sc.arrays()
    b = hiprt.Bvh(sc)
    assert len(b.index) > sc.n_tris      # clipping happened
    parent_box = {}
    for nid, nd in enumerate(b.nodes):
        if nd["link_r"] >= 0:
            parent_box[int(nd["link_l"])] = nd["box_l"]
            parent_box[int(nd["link_r"])] = nd["box_r"]
    boxes_of = {}
    for nid, nd in enumerate(b.nodes):
        if nd["link_r"] < 0:
            for s in range(-nd["link_l"], -nd["link_l"] - nd["link_r"]):
                boxes_of.setdefault(int(b.index[s]), []).append(parent_box[nid])
    rng = np.random.default_rng(0)
    dup = [t for t, bx in boxes_of.items() if len(bx) > 1][:200]
    assert dup
    for t in dup:
        A, B, C = pos[idx[t]].astype(np.float64)
        for _ in range(20):
            r1, r2 = rng.random(2)
            s = np.sqrt(r1)
            p = (1 - s) * A + s * (1 - r2) * B + s * r2 * C
            assert any(np.all(bx[0] - 1e-5 <= p) and np.all(p <= bx[1] + 1e-5) for bx in boxes_of[t])


# ---------------------------------------------------------------- ray generation

def test_config_camera_rays_and_depth_fit_the_kernels():
    for cfg in (1, 2, 3):
        sc = hiprt.HostScene.generate(cfg)
        b = hiprt.Bvh(sc)
        # persistent kernels keep 32 stack entries; the plain kernel 63
        assert b.depth <= 32, (cfg, b.depth)


def test_bvh_build_ends_on_large_coordinates():
    """Triangles out to 2^49 along x: float rounding collapses most of them to
    flat boxes (a zero split threshold) and stops midpoint splits from
    shrinking the far ones.  The build must end and give a valid tree."""
    n = 50
    c = np.stack([2.0 ** np.arange(n), np.zeros(n), np.zeros(n)], 1)
    pos = (c[:, None, :] + np.random.default_rng(1).normal(0, 0.05, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    sc = hiprt.HostScene.from_arrays(pos, idx)
    for max_tris in (1, 4):
        bvh = hiprt.Bvh(sc, max_tris=max_tris)
        assert bvh.validate()[0] == 0
        assert np.array_equal(np.unique(bvh.index), np.arange(n))
