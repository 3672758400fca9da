"""The script front end (illum-rt_amd/rtgi-hip): RTGI's command protocol on
the hip platform.  CPU tests check the error behaviour the reference's
interaction.cpp has; GPU tests run whole scripts and compare the written
images with the oracle / the library."""
import os
import subprocess

import numpy as np
import pytest

import hiprt
import oracle_py as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "illum-rt_amd", "rtgi-hip")
FLT_MAX32 = np.float32(3.4028234663852886e38)


def run_script(text, tmp_path, timeout=120):
    p = tmp_path / "script.rtgi"
    p.write_text(text)
    r = subprocess.run([DRIVER, str(p)], capture_output=True, text=True, timeout=timeout, cwd=tmp_path)
    return r


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, 3)


def test_driver_errors_follow_the_reference(tmp_path):
    r = run_script("echo hi\nplatform cuda\nraytracer while-while\nrun\nresolution 0 5\nsppx x\nbogus\n"
                   "generate 9\nbvh type sah\ntracer-option stack 3\n# comment\n// comment\n", tmp_path)
    assert r.returncode == 0
    err = r.stderr.splitlines()
    # "command <n> (<name>): ..." with n counting every script line (interaction.cpp:96, 226)
    assert "command 2 (platform): There is no platform called 'cuda'" in err
    assert any(e.startswith("command 3 (raytracer): No platform selected") for e in err)
    assert "command 4 (run): Invalid platform" in err
    assert "command 5 (resolution): Resolution must be positive" in err
    assert "command 6 (sppx): Syntax error, requires exactly one positive integral value" in err
    assert "command 7 (bogus): Unknown command" in err
    assert any(e.startswith("command 10 (tracer-option): No platform selected") for e in err)
    assert r.stdout.strip() == "hi"
    assert len(err) == 9


def test_driver_blocks_and_material_commands(tmp_path):
    """The reference's "with <feature> {" blocks (interaction.cpp:263-297):
    gl, cuda and optix are not built here, so their blocks are skipped up to
    the closing "}"; "with hip {" runs unless hip was disabled.  "material"
    (interaction.cpp:624-706): list, select by name, albedo/emissive, the
    reference's error messages."""
    (tmp_path / "lib.mtl").write_text("newmtl floor\nKd 0.5 0.5 0.5\nnewmtl lamp\nKd 0.5 0.5 0.5\n")
    (tmp_path / "s.obj").write_text("mtllib lib.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl floor\nf 1 2 3\n")
    r = run_script("with cuda {\n\tplatform cuda\n\techo never\n}\necho after\nwith hip {\necho inside\n}\n"
                   "disable hip\nwith hip {\necho never\n}\nwith vulkan {\n"
                   "load s.obj\nmaterial list\nmaterial albedo 1 1 1\nmaterial select nope\nmaterial select lamp\n"
                   "material emissive 5 4 3\nmaterial roughness 0.3\nmaterial show\nmaterial frobnicate\n"
                   "default-brdf layered-gtr2\n", tmp_path)
    assert r.returncode == 0
    out = r.stdout.splitlines()
    assert out[:5] == ["Skipping a block due to missing cuda", "after", "inside", "DISABLING hip",
                       "Skipping a block due to missing hip"]
    assert "never" not in out
    assert out[5:8] == ["DefaultMaterial", "floor", "lamp"]
    assert "emissive   5 4 3" in out and "roughness  0.3" in out
    err = r.stderr.splitlines()
    # errors inside "material" name the subcommand (interaction.cpp:632); skipped lines count too
    assert err == ["command 13 (with): Syntax error: expected one of gl, cuda, optix, hip",
                   "command 16 (albedo): No material selected",
                   "command 17 (select): No material called 'nope'",
                   "command 22 (frobnicate): Unknown subcommand"], err
    assert out[-1] == "Replacing default brdfs that has already been applied to some materials."


SCRIPT_HEAD = """platform hip
generate {cfg}
resolution {w} {h}
raytracer {tracer}
add-scene-step build accel struct
commit
"""


@pytest.mark.gpu
@pytest.mark.parametrize("tracer", ["while-while", "default", "simple", "persistent-if-if"])
def test_primary_wf_image_matches_oracle(tmp_path, torch_cuda, tracer):
    print(f"case: rtgi-hip primary-wf, 8k random triangles, 160x120, raytracer {tracer}", flush=True)
    w, h = 160, 120
    r = run_script(SCRIPT_HEAD.format(cfg=1, w=w, h=h, tracer=tracer) +
                   "algo primary-wf\noutfile p.pfm\nrun\nstats print\n", tmp_path)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    img = read_pfm(tmp_path / "p.pfm")
    sc = hiprt.HostScene.generate(1)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    cam = sc.cam.copy()
    cam[10], cam[11] = w, h
    rays = hiprt.camera_rays(cam)
    t, *_ = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    want = np.where(t != np.float32(3.4028234663852886e38), np.float32(0.8), np.float32(0)).reshape(h, w)
    assert np.array_equal(img[..., 0], want) and np.array_equal(img[..., 2], want)
    assert "find closest hit" in r.stdout


@pytest.mark.gpu
def test_tracer_options_keep_the_image(tmp_path, torch_cuda):
    print("case: rtgi-hip primary-wf with tracer options, 8k random triangles, 160x120", flush=True)
    w, h = 160, 120
    head = SCRIPT_HEAD.format(cfg=1, w=w, h=h, tracer="dynamic-while-while")
    r0 = run_script(head + "algo primary-wf\noutfile a.pfm\nrun\n", tmp_path)
    r1 = run_script(head + "tracer-option stack 4\ntracer-option chunks 8\ntracer-option refill 56\n"
                    "tracer-option stack 9\ntracer-option speed 1\nalgo primary-wf\noutfile b.pfm\nrun\n", tmp_path)
    assert r0.returncode == 0 and not r0.stderr.strip(), r0.stderr
    err = r1.stderr.splitlines()
    # the platform's own messages have no command number, as rt/cuda/platform.cpp:18 prints them
    assert err == ["command (tracer-option): Parameter error, stack configuration must be 0..4",
                   "command (tracer-option): Unknown tracer option speed (refill, blocks, stack, chunks, ray-order, samples, "
                   "park, regen)"], err
    # primary-wf shades hits with the albedo, so the images agree unless two hits tie in t
    a, b = read_pfm(tmp_path / "a.pfm"), read_pfm(tmp_path / "b.pfm")
    assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["uniform", "cosine", "light"])
def test_direct_wf_matches_oracle(tmp_path, torch_cuda, mode):
    """direct-wf with "is <mode>" (uniform is the default) on the interior hall
    with its emissive arches and a point light, 2 samples: the written image
    is the oracle's framebuffer divided by its weights.  Sample k of the
    sampling step draws from seed k (1, 2); light mode is bit-exact, the
    others may differ where a cosf/sinf ulp turns a sample onto another
    triangle (tests/test_gpu_direct.py)."""
    print(f"case: rtgi-hip direct-wf, hall 192x108, is {mode}, 2 spp", flush=True)
    w, h = 192, 108
    r = run_script(SCRIPT_HEAD.format(cfg=2, w=w, h=h, tracer="persistent-while-while") +
                   "pointlight replace pos 20 13.5 7.5 col 40 40 40\nsppx 2\nalgo direct-wf\n" +
                   ("" if mode == "uniform" else f"is {mode}\n") + "outfile d.pfm\nrun\n", tmp_path)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    img = read_pfm(tmp_path / "d.pfm")
    sc = hiprt.HostScene.generate(2)
    pos, idx, mat = sc.arrays()
    bvh = hiprt.Bvh(sc)
    cam = sc.cam.copy()
    cam[10], cam[11] = w, h
    fb = orc.direct(pos, idx, mat, sc.materials(), bvh.nodes, bvh.root, bvh.index, cam, mode,
                    np.array([1, 2], np.uint64), np.array([[20, 13.5, 7.5, 40, 40, 40]], np.float32))
    want = fb[..., :3] / fb[..., 3:4]
    differ = (img.view(np.uint32) != want.view(np.uint32)).any(-1)
    scale = np.maximum(np.abs(want).max(-1), 1e-30)
    far = np.abs(img - want).max(-1) > 1e-5 * scale
    print({"pixels_differ": int(differ.sum()), "pixels_beyond_rel": int(far.sum())})
    assert want.sum() > 0
    if mode == "light":
        assert not differ.any()
    else:   # last-bit cosine differences everywhere, another triangle rarely
        assert far.sum() <= max(3, 4e-4 * far.size)


@pytest.mark.gpu
@pytest.mark.parametrize("options,samples", [("", 1), ("tracer-option ray-order 1\ntracer-option stack 3\n", 1),
                                             ("tracer-option stack 4\ntracer-option samples 2\n", 2),
                                             ("tracer-option samples 2\ntracer-option regen 1\ntracer-option park 3\n",
                                              2)])
def test_pt_wf_matches_library_path_tracer(tmp_path, torch_cuda, options, samples):
    print(f"case: rtgi-hip pt-wf, hall 128x72, 2 spp vs hiprt.PathTracer, options {options!r}", flush=True)
    w, h = 128, 72
    r = run_script(SCRIPT_HEAD.format(cfg=3, w=w, h=h, tracer="persistent-while-while") + options +
                   "sppx 2\nalgo pt-wf\noutfile pt.pfm\nrun\n", tmp_path, timeout=300)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    img = read_pfm(tmp_path / "pt.pfm")
    sc = hiprt.HostScene.generate(3)
    bvh = hiprt.Bvh(sc)
    ctx = hiprt.Context(0)
    ds = hiprt.DeviceScene(ctx, sc, bvh)
    cam = sc.cam.copy()
    cam[10], cam[11] = w, h
    # with k samples per render call, calls 0 .. 2/k - 1 trace frames 0 and 1
    pt = hiprt.PathTracer(ds, cam, seed=1, batched=True, samples=samples, tracer="persistent-while-while")
    for f in range(2 // samples):
        pt.render(f)
    want = pt.image() / np.float32(2)
    pt.close()
    ds.close()
    ctx.close()
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_pt_wf_path_commands(tmp_path, torch_cuda):
    """simple_pt's "path len n" / "path rr-start n" (gi/pt.cpp:121-154) on
    pt-wf: the reference's loop traces rays 0 .. n-1 and, shading the hit of
    ray i > rr-start, plays roulette on ray i + 1, so "len 7" traces levels
    0..6 and "rr-start 1" subjects levels 3 on; the image equals the
    library's path tracer with max_depth 6 and rr_start 3.  Bad values print
    the reference's messages and change nothing."""
    print("case: rtgi-hip pt-wf, hall 128x72, 2 spp, path len 7 / rr-start 1 vs hiprt.PathTracer", flush=True)
    w, h = 128, 72
    r = run_script(SCRIPT_HEAD.format(cfg=3, w=w, h=h, tracer="persistent-while-while") +
                   "sppx 2\nalgo pt-wf\npath len 7\npath rr-start 1\npath bounce brdf\npath len 0\n"
                   "path rr-start -2\npath bounce uniform\npath mis on\noutfile pt.pfm\nrun\n", tmp_path, timeout=300)
    assert r.returncode == 0
    assert r.stderr.splitlines() == ["error in path len: expected a positive integer, got 0",
                                     "error in russian roulette offset: expected a positive integer > 0, got -2",
                                     "error: pt-wf samples bounces by the BRDF (cosine) only",
                                     "unknown subcommand to path: 'mis'"], r.stderr
    img = read_pfm(tmp_path / "pt.pfm")
    sc = hiprt.HostScene.generate(3)
    bvh = hiprt.Bvh(sc)
    ctx = hiprt.Context(0)
    ds = hiprt.DeviceScene(ctx, sc, bvh)
    cam = sc.cam.copy()
    cam[10], cam[11] = w, h
    pt = hiprt.PathTracer(ds, cam, max_depth=6, rr_start=3, seed=1, batched=True, tracer="persistent-while-while")
    for f in range(2):
        pt.render(f)
    want = pt.image() / np.float32(2)
    levels = pt.ray_totals()
    pt.close()
    ds.close()
    ctx.close()
    assert len(levels) == 7 and levels[6] > 0
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_incoherence_rays_in_the_camera_step(tmp_path, torch_cuda):
    """"incoherence r1 r2" (rt/cuda/base.cpp:120-133): the camera step of
    primary-wf then traces rays between two balls placed along the scene's
    largest extent (rt/cuda/rni.cpp:36-89), the same rays every sample
    (seed 1234); the image is the albedo at the oracle's hits on the oracle's
    rays, up to the rays the device's libm ulps move across an edge.  The
    platform's own messages follow base.cpp; "0 0" keeps the setting."""
    print("case: rtgi-hip primary-wf, hall 160x90, incoherence 0.5 0.5", flush=True)
    w, h = 160, 90
    r = run_script(SCRIPT_HEAD.format(cfg=2, w=w, h=h, tracer="while-while") +
                   "incoherence 1\nincoherence -1 2\nincoherence 0.5 0.5\nincoherence 0 0\nsppx 2\n"
                   "algo primary-wf\noutfile i.pfm\nrun\n", tmp_path)
    assert r.returncode == 0
    assert r.stderr.splitlines() == [
        'command (incoherence): Syntax error, "incoherence" requires exactly two positive float values',
        'command (incoherence): Parameter error, "incoherence" requires exactly two positive float values'], r.stderr
    img = read_pfm(tmp_path / "i.pfm").reshape(-1, 3)
    sc = hiprt.HostScene.generate(2)
    pos, idx, mat = sc.arrays()
    s1, s2, r_max = hiprt.incoherence_spheres(pos)
    rays = orc.incoherent_rays(s1, s2, 0.5, 0.5, r_max, 1234, w * h)
    bvh = hiprt.Bvh(sc)
    t, _, _, tid = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    hit = t != FLT_MAX32
    want = np.zeros((w * h, 3), np.float32)
    want[hit] = sc.materials()[mat[tid[hit]], :3]
    assert hit.mean() > 0.9 and len(np.unique(mat[tid[hit]])) > 2
    same = (img == want).all(axis=1).mean()
    assert same >= 0.995, same


@pytest.mark.gpu
def test_pt_wf_samples_must_divide_sppx(tmp_path, torch_cuda):
    print("case: rtgi-hip pt-wf, sppx 3 with tracer-option samples 2", flush=True)
    r = run_script(SCRIPT_HEAD.format(cfg=1, w=32, h=24, tracer="dynamic-while-while") +
                   "tracer-option samples 2\ntracer-option samples 65\nsppx 3\nalgo pt-wf\noutfile pt.pfm\nrun\n",
                   tmp_path)
    err = r.stderr.splitlines()
    assert "command (tracer-option): Parameter error, samples must be 1..64" in err
    assert "command 12 (run): pt-wf: sppx 3 is not a multiple of tracer-option samples 2" in err


@pytest.mark.gpu
def test_unknown_tracer_and_stale_bvh(tmp_path, torch_cuda):
    r = run_script("platform hip\nraytracer if-else\ngenerate 1\nalgo primary-wf\nrun\n", tmp_path)
    err = r.stderr.splitlines()
    assert "command (raytracer): There is no such hip ray tracer variant" in err
    assert any(e.startswith("command 5 (run): We have to have a scene loaded") for e in err)


CUBE_OBJ = ("mtllib lib.mtl\n" + "".join(f"v {x} {y} {z}\n" for x in (0, 1) for y in (0, 1) for z in (0, 1)) +
            "usemtl red\nf 1 2 4 3\nf 5 7 8 6\nf 1 5 6 2\nf 3 4 8 7\nf 1 3 7 5\nf 2 6 8 4\n")


def test_load_accumulates_and_finds_models(tmp_path):
    """"load" appends models (interaction.cpp:395-402): material names
    accumulate, a named model's as "<name>/<material>"; relative paths are
    found under "modelpath" entries and beside the script; "modeltrafo" and
    "modelmatrix" report the reference's syntax errors."""
    models = tmp_path / "models"
    models.mkdir()
    (models / "lib.mtl").write_text("newmtl floor\nKd 0.5 0.5 0.5\nnewmtl red\nKd 0.9 0.2 0.1\n")
    (models / "cube.obj").write_text(CUBE_OBJ)
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "near.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    script = tmp_path / "sub" / "s.rtgi"
    script.write_text("load models/cube.obj\nmodelpath models\nload cube.obj box\nload near.obj\nmaterial list\n"
                      "modeltrafo spin 1\nmodeltrafo rotate w 3\nmodeltrafo scale 1 2\nmodelmatrix 1 2 3\n"
                      "load missing.obj\n")
    r = subprocess.run([DRIVER, str(script)], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 0
    assert r.stdout.splitlines() == ["DefaultMaterial", "floor", "red", "box/DefaultMaterial", "box/floor",
                                     "box/red", "DefaultMaterial"]
    err = r.stderr.splitlines()
    assert err[:4] == ["command 6 (modeltrafo): unsupported trafo: spin",
                       "command 7 (modeltrafo): Invalid axis of rotation; w",
                       "command 8 (modeltrafo): Too many components in scale transformation",
                       "command 9 (modelmatrix): Syntax error in model matrix"], err
    assert len(err) == 5 and err[4].startswith("command 10 (load): cannot open missing.obj")


def test_camera_sky_and_host_commands(tmp_path):
    """"up" sets the scene's up vector once and the camera's unless "camup"
    did; at/look/up/camup without arguments print the vector
    (interaction.cpp:301-350).  skylight, preview-offset, omp and denoise are
    checked with the reference's messages (interaction.cpp:573-605, 754-777,
    819-828); the BRDF names are new_brdf's (libgi/material.cpp:326-370)."""
    (tmp_path / "sky.floats").write_bytes(b"\0" * 16)
    r = run_script("up 0 0 1\nup 0 1 0\ncamup\ncamup 1 0 0\ncamup\nup\nat 1 2 3 4\nat\nskylight scale 2\n"
                   "skylight load sky.floats 30\nskylight scale -1\nskylight load none.floats 3\nskylight fly\n"
                   "preview-offset 0\nomp maybe\nomp on\ndenoise on\ndenoise soon\ndefault-brdf phong\n"
                   "default-brdf fancy\nskylight scale 4\npreview-offset 8\ndenoise off\n", tmp_path)
    assert r.returncode == 0
    assert r.stdout.splitlines() == ["0 0 1", "1 0 0", "0 0 1", "0 0 0",
                                     "Denoising not available: OpenImageDenoise library not found"]
    assert r.stderr.splitlines() == [
        "command 2 (up): Cannot set scene up vector twice, did you mean camup?",
        "command 7 (at): Syntax error, requires 3 numerical components",
        "command 9 (skylight): There is no skylight to scale",
        "command 11 (skylight): Intensity scale has to be a value > 0",
        "command 12 (skylight): Cannot open file 'none.floats' for hdr floats texture.",
        "command 13 (skylight): No such skylight subcommand",
        "command 14 (preview-offset): Thge preview offset needs to be larger than zero",
        "command 15 (omp): Syntax error: expected 'on' or 'off'",
        "command 18 (denoise): Invalid option for denoise, valid are on/off",
        "command 20 (default-brdf): No such brdf defined: fancy"]


@pytest.mark.gpu
def test_load_under_modeltrafo_matches_oracle(tmp_path, torch_cuda):
    """Two models, the second under "modeltrafo translate / rotate / scale"
    (right-multiplied, the rotation angle taken as radians as the
    reference's matrix library takes it): primary-wf's albedo image equals
    the oracle's closest hits on the scene hiprt_scene_add_obj builds with
    the same matrix -- bit for bit for translate and scale, whose products
    are exact, and with a rotation up to a few silhouette pixels, where the
    driver's float matrix and numpy's rounded one differ in the last bit."""
    print("case: rtgi-hip primary-wf, floor + cube under modeltrafo, 128x96", flush=True)
    models = tmp_path / "models"
    models.mkdir()
    (models / "lib.mtl").write_text("newmtl floor\nKd 0.5 0.5 0.5\nnewmtl red\nKd 0.9 0.2 0.1\n")
    (models / "floor.obj").write_text("mtllib lib.mtl\nv -4 0 -4\nv 4 0 -4\nv 4 0 4\nv -4 0 4\nusemtl floor\n"
                                      "f 1 4 3 2\n")
    (models / "cube.obj").write_text(CUBE_OBJ)
    w, h = 128, 96
    cam = np.array([0, 6, -7, 0, -0.7, 1, 0, 1, 0, 60, w, h], np.float32)
    a = 1.25

    def T(x, y, z):
        m = np.eye(4)
        m[:3, 3] = (x, y, z)
        return m

    def S(x, y, z):
        return np.diag([x, y, z, 1.0])

    Ry = np.eye(4)
    Ry[[0, 0, 2, 2], [0, 2, 0, 2]] = [np.cos(a), np.sin(a), -np.sin(a), np.cos(a)]
    cases = [("translate 1 0 0.5\nmodeltrafo scale 2 0.5 2\n", T(1, 0, 0.5) @ S(2, 0.5, 2), 1.0),
             (f"translate -1 0 1\nmodeltrafo rotate y {a}\nmodeltrafo scale 1.5 1 1.5\n",
              T(-1, 0, 1) @ Ry @ S(1.5, 1, 1.5), 0.995)]
    for k, (trafo, M, bar) in enumerate(cases):
        r = run_script("platform hip\nmodelpath models\nload floor.obj\nmodeltrafo " + trafo +
                       "load cube.obj box\nat 0 6 -7\nlook 0 -0.7 1\nup 0 1 0\nfovy 60\n"
                       f"resolution {w} {h}\nraytracer while-while\nadd-scene-step build accel struct\ncommit\n"
                       f"algo primary-wf\noutfile p{k}.pfm\nrun\n", tmp_path)
        assert r.returncode == 0 and not r.stderr.strip(), r.stderr
        img = read_pfm(tmp_path / f"p{k}.pfm").reshape(-1, 3)
        sc = hiprt.HostScene.add_obj(models / "floor.obj")
        hiprt.HostScene.add_obj(models / "cube.obj", name="box", trafo=M.astype(np.float32), scene=sc)
        pos, idx, mat = sc.arrays()
        table = sc.materials()
        bvh = hiprt.Bvh(sc)
        t, _, _, tid = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, hiprt.camera_rays(cam))
        hit = t != np.float32(3.4028234663852886e38)
        want = np.zeros((w * h, 3), np.float32)
        want[hit] = table[mat[tid[hit]], :3]
        assert (table[mat[tid[hit]], 0] > 0.5).any()   # the cube is in view
        same = (img == want).all(axis=1).mean()
        assert same >= bar, (k, same)


def _quarter_turns():
    """Rz(pi/4) Ry(pi/4) Rx(pi/4) in float32, R[column][row]."""
    f = np.float32
    a = f(np.pi) / f(4)
    c, s = f(np.cos(a)), f(np.sin(a))
    Rx = np.array([[1, 0, 0], [0, c, s], [0, -s, c]], f)      # [column][row]
    Ry = np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]], f)
    Rz = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]], f)

    def mul(A, B):   # (A B)[col][r] = sum_k A[k][r] B[col][k]
        return np.array([[(A[0][r] * B[col][0] + A[1][r] * B[col][1]) + A[2][r] * B[col][2] for r in range(3)]
                         for col in range(3)], f)

    return mul(mul(Rz, Ry), Rx)


@pytest.mark.gpu
def test_rotate_scene_steps_match_oracle(tmp_path, torch_cuda):
    """The reference's example scene steps (rt/cuda/preprocessing.cpp:41-72):
    "rotate scene" turns the vertices by Rz Ry Rx (pi/4 each) before "build
    accel struct"; "rotate scene keeping original data" then "drop scene
    view" traces the turned scene and restores the original for the next
    commit, which turns it again.  primary-wf's image equals the oracle's
    hits on the turned cube (up to silhouette pixels, where numpy's float
    cosine may differ from the host library's in the last bit); dropping a
    view that was never made is an error."""
    print("case: rtgi-hip primary-wf, cube turned by the rotate scene steps, 96x96", flush=True)
    (tmp_path / "cube.obj").write_text(
        "".join(f"v {x} {y} {z}\n" for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)) +
        "f 1 2 4 3\nf 5 7 8 6\nf 1 5 6 2\nf 3 4 8 7\nf 1 3 7 5\nf 2 6 8 4\n")
    w = h = 96
    cam = np.array([0, 0, -3, 0, 0, 1, 0, 1, 0, 40, w, h], np.float32)
    head = (f"platform hip\nload cube.obj\nat 0 0 -3\nlook 0 0 1\nup 0 1 0\nfovy 40\nresolution {w} {h}\n"
            "raytracer while-while\n")
    tail = "algo primary-wf\noutfile {}\nrun\n"
    r = run_script(head + "add-scene-step rotate scene\nadd-scene-step build accel struct\ncommit\n" +
                   tail.format("a.pfm"), tmp_path)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    r = run_script(head + "add-scene-step rotate scene keeping original data\nadd-scene-step build accel struct\n"
                   "add-scene-step drop scene view\ncommit\ncommit\n" + tail.format("b.pfm"), tmp_path)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    r = run_script(head + "add-scene-step drop scene view\ncommit\n", tmp_path)
    assert "command 10 (commit): drop scene view: there is no scene view to drop" in r.stderr.splitlines(), r.stderr
    sc = hiprt.HostScene.load_obj(tmp_path / "cube.obj")
    pos, idx, mat = sc.arrays()
    R = _quarter_turns()
    turned = np.stack([(pos[:, 0] * R[0][r] + pos[:, 1] * R[1][r]) + pos[:, 2] * R[2][r] for r in range(3)], 1)
    sc.set_positions(turned.astype(np.float32))
    bvh = hiprt.Bvh(sc)
    t, _, _, tid = orc.trace_closest(turned.astype(np.float32), idx, bvh.nodes, bvh.root, bvh.index,
                                     hiprt.camera_rays(cam))
    hit = t != np.float32(3.4028234663852886e38)
    want = np.zeros((w * h, 3), np.float32)
    want[hit] = sc.materials()[mat[tid[hit]], :3]
    bvh0 = hiprt.Bvh(hiprt.HostScene.load_obj(tmp_path / "cube.obj"))   # the unturned cube
    t0, *_ = orc.trace_closest(pos, idx, bvh0.nodes, bvh0.root, bvh0.index, hiprt.camera_rays(cam))
    plain = t0 != np.float32(3.4028234663852886e38)
    assert (hit != plain).sum() > 100   # the turn changes the silhouette
    for name in ("a.pfm", "b.pfm"):
        img = read_pfm(tmp_path / name).reshape(-1, 3)
        same = (img == want).all(axis=1).mean()
        assert same >= 0.995, (name, same)


@pytest.mark.gpu
def test_material_commands_light_an_obj_scene(tmp_path, torch_cuda):
    """An OBJ scene whose lamp material has no emission in its library gets
    one by "material select lamp" / "material emissive" before commit, as
    RTGI's scripts light sibenik's windows; direct-wf in light mode then
    equals the oracle over the edited material table, bit for bit."""
    print("case: rtgi-hip direct-wf light mode on an OBJ floor under a lamp, material emissive set by script",
          flush=True)
    (tmp_path / "lib.mtl").write_text("newmtl floor\nKd 0.6 0.6 0.6\nnewmtl lamp\nKd 0.5 0.5 0.5\n")
    (tmp_path / "s.obj").write_text(
        "mtllib lib.mtl\nv -2 0 -2\nv 2 0 -2\nv 2 0 2\nv -2 0 2\nv -0.5 1 -0.5\nv 0.5 1 -0.5\nv -0.5 1 0.5\n"
        "usemtl floor\nf 1 3 2\nf 1 4 3\nusemtl lamp\nf 5 6 7\n")
    w, h = 64, 48
    r = run_script("platform hip\nload s.obj\nat 0.1 0.7 0.05\nlook 0 -1 0\nup 0 0 1\nfovy 60\n"
                   f"resolution {w} {h}\nraytracer default\nmaterial select lamp\nmaterial emissive 5 4 3\n"
                   "add-scene-step build accel struct\ncommit\nsppx 2\nalgo direct-wf\nis light\n"
                   "outfile d.pfm\nrun\n", tmp_path)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr
    img = read_pfm(tmp_path / "d.pfm")
    sc = hiprt.HostScene.load_obj(tmp_path / "s.obj")
    sc.set_material(2, emissive=[5, 4, 3])
    pos, idx, mat = sc.arrays()
    bvh = hiprt.Bvh(sc)
    cam = np.array([0.1, 0.7, 0.05, 0, -1, 0, 0, 0, 1, 60, w, h], np.float32)
    fb = orc.direct(pos, idx, mat, sc.materials(), bvh.nodes, bvh.root, bvh.index, cam, "light",
                    np.array([1, 2], np.uint64), None)
    want = fb[..., :3] / fb[..., 3:4]
    assert want.sum() > 0
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))
