"""Host side of the boundary (CPU): the C-ABI library loads and exports every
symbol include/orion_amd.h declares; the .ori loader's resolved records equal
the reference's own transform / camera / light code (oracle/_ref); writer ->
loader round trips are bit-exact; errors come back as status codes."""
import re
from pathlib import Path

import numpy as np
import pytest

import orion_amd as oa

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_header_symbols():
    import ctypes
    hdr = (ROOT / "include" / "orion_amd.h").read_text()
    names = set(re.findall(r"\b(orion_[a-z_]+)\s*\(", hdr))
    lib = ctypes.CDLL(str(oa.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(oa.exported_symbols()) <= names


def test_scene_sizes():
    assert oa.HostScene.generate("cornell").n_tris == 36
    assert oa.HostScene.generate("caustics").n_tris == 81920 + 2
    hs = oa.HostScene.generate("icosphere", params=[4, 64, 64])
    assert hs.n_tris == 20 * 4 ** 4 + 4


def test_full_size_icosphere():
    hs = oa.HostScene.generate("icosphere")
    assert hs.n_tris == 1310724 and hs.width == 1024 and hs.info.spp == 64 and hs.info.max_depth == 8
    n = hs.tri_pn.reshape(-1, 18)[:, 9:]
    assert np.isfinite(hs.tri_pn).all() and np.allclose(np.linalg.norm(n.reshape(-1, 3), axis=1), 1, atol=1e-5)


def test_create_transform_matches_reference(ref):
    rng = np.random.default_rng(3)
    chains = [
        [("translate", "1,2,3")],
        [("scale", "2,0.5,1.5"), ("rotate", "0,1,0,30")],
        [("rotate", "1,1,0,45"), ("translate", "-0.5,0.25,4"), ("scale", "1,1,-1")],
        [("lookat", "0,1,5, 0,0,0, 0,1,0")],
        [("translate", "1,0,0"), ("store", "1"), ("rotate", "0,0,1,90"), ("reuse", "1")],
    ]
    for _ in range(10):
        chains.append([("rotate", ",".join(f"{x:.4f}" for x in rng.normal(size=3)) + f",{rng.uniform(-180, 180):.3f}"),
                       ("translate", ",".join(f"{x:.3f}" for x in rng.normal(size=3)))])
    for ch in chains:
        m, mi = oa.create_transform(ch)
        rm, rmi = ref.create_transform(ch)
        np.testing.assert_array_equal(m.view(np.int32), rm.view(np.int32))
        np.testing.assert_array_equal(mi.view(np.int32), rmi.view(np.int32))


def test_camera_record_matches_reference(ref):
    rng = np.random.default_rng(9)
    for _ in range(50):
        pos = rng.normal(size=3) * 4
        look = rng.normal(size=3)
        up = np.array([0, 1, 0]) + 0.1 * rng.normal(size=3)
        la = np.concatenate([pos, look, up]).astype(np.float32)
        fov = float(rng.uniform(10, 90))
        W, H = int(rng.integers(16, 2048)), int(rng.integers(16, 2048))
        rec = oa.camera_record(la, fov, W, H)
        m, ta, ray = ref.camera(la, fov, W, H)
        # origin = the transform applied to (0,0,0), including the w divide (transform.h:118-131)
        np.testing.assert_array_equal(rec[:3].view(np.int32), ray[:3].view(np.int32))
        np.testing.assert_array_equal(rec[3:12].view(np.int32), m[:3, :3].reshape(-1).view(np.int32))
        assert rec[12] == ta[0] and rec[13] == ta[1]


@pytest.mark.parametrize("scene,params,limit", [("cornell", None, 1000), ("icosphere", [3, 64, 48], 100),
                                                ("caustics", [2, 40, 30], 10 ** 9),
                                                ("materials", [32, 32], 1000), ("sky", [32, 32], 1000)])
def test_ori_round_trip(tmp_path, scene, params, limit):
    hs = oa.HostScene.generate(scene, params=params)
    path = tmp_path / "scene.ori"
    hs.write_ori(path, inline_limit=limit)
    back = oa.HostScene.load_ori(path)
    for name in ("tri_pn", "tri_light", "lights", "camera"):
        a, b = getattr(hs, name), getattr(back, name)
        assert a.shape == b.shape, name
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32), err_msg=name)
    # the loader makes one material per Model (parser.cpp:95-96), so indices are
    # renumbered; the resolved material of every triangle is unchanged
    ma = hs.mats.reshape(-1, oa.MAT_FLOATS)[hs.tri_mat]
    mb = back.mats.reshape(-1, oa.MAT_FLOATS)[back.tri_mat]
    np.testing.assert_array_equal(ma.view(np.int32), mb.view(np.int32))
    assert back.info.integrator == hs.info.integrator and back.info.max_depth == hs.info.max_depth
    assert len(back.env_images) == len(hs.env_images)
    for a, b in zip(hs.env_images, back.env_images):
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))


def test_loader_schema_features(tmp_path):
    """Transforms on models, named materials, metal presets, one-normal meshHandy."""
    (tmp_path / "s.ori").write_text("""<?xml version="1.0"?>
<Scene>
  <Resource filename="./"/>
  <Camera type="pinhole"><Transform><Param lookat="0,0,5, 0,0,0, 0,1,0"/></Transform>
    <Param fov="40"/><Param filmsize="32,24"/></Camera>
  <!-- a comment -->
  <Model type="meshHandy">
    <Transform><Param scale="2,2,2"/><Param translate="0,1,0"/></Transform>
    <Param p="0,0,0, 1,0,0, 0,1,0"/><Param n="0,0,1"/>
    <Material type="metal" name="gold1"><Param metalName="gold"/>
      <roughness><Param type="float"/><Param value="0.2"/></roughness></Material>
  </Model>
  <Model type="meshHandy">
    <Param p="0,0,1, 1,0,1, 0,1,1"/><Param n="0,0,1, 0,0,1, 0,0,1"/>
    <Material named="gold1"/>
    <AreaLight><Param Le="4"/><Param nLightSamples="1"/></AreaLight>
  </Model>
  <Light type="point"><Param I="1,2,3"/><Param from="0,4,0"/></Light>
  <Integrator type="path"><Param maxDepth="5"/><Param nSamplesPerPixel="4"/></Integrator>
  <Sampler type="sobol"/>
</Scene>""")
    hs = oa.HostScene.load_ori(tmp_path / "s.ori")
    assert hs.n_tris == 2 and hs.n_mats == 2 and hs.n_lights == 2
    p = hs.tri_pn.reshape(-1, 18)
    np.testing.assert_array_equal(p[0, :9], [0, 1, 0, 2, 1, 0, 0, 3, 0])
    m = hs.mats.reshape(-1, oa.MAT_FLOATS)
    np.testing.assert_array_equal(m[0], m[1])
    assert m[0][0] == 1 and m[0][7] == np.float32(0.2)
    L = hs.lights.reshape(-1, 32)
    assert L[0][0] == 0 and L[0][1] == 1 and (L[0][2:5] == 4).all()
    assert L[1][0] == 1 and (L[1][2:5] == [1, 2, 3]).all() and (L[1][5:8] == [0, 4, 0]).all()
    assert list(hs.tri_light) == [-1, 0]
    assert hs.info.spp == 4 and hs.info.max_depth == 5 and hs.width == 32 and hs.height == 24


def test_loader_materials(tmp_path, ref):
    """mirror / plastic / uber / Oren-Nayar matte records (parser.cpp:262-421)."""
    tri = "<Param p='0,0,0, 1,0,0, 0,1,0'/><Param n='0,0,1'/>"
    def c(n, v):
        return f"<{n}><Param type='constant'/><Param color='{v}'/></{n}>"
    def f(n, v):
        return f"<{n}><Param type='float'/><Param value='{v}'/></{n}>"
    mats = [
        "<Material type='mirror'>" + c("Kr", "0.9,0.8,0.7") + "</Material>",
        "<Material type='plastic'>" + c("Kd", "0.1,0.2,0.3") + f("Ks", "0.4") + f("roughness", "0.12") +
        "</Material>",
        "<Material type='uber'>" + c("Kd", "0.5,0.4,0.3") + c("Ks", "0.2,0.2,0.2") + c("Kr", "0.1,0.1,0.1") +
        c("Kt", "0.3,0.3,0.3") + c("opacity", "0.6,0.7,0.8") + f("roughness", "0.2") + f("eta", "1.45") +
        "</Material>",
        "<Material type='matte'>" + c("Kd", "0.5") + f("sigma", "20") + "</Material>",
        # a constant (non-float) sigma fails the FloatTexture cast: Lambert
        "<Material type='matte'>" + c("Kd", "0.5") + c("sigma", "20,20,20") + "</Material>",
    ]
    body = ("<Scene><Camera type='pinhole'><Param fov='40'/><Param filmsize='4,4'/></Camera>" +
            "".join(f"<Model type='meshHandy'>{tri}{m}</Model>" for m in mats) +
            "<Integrator type='path'><Param maxDepth='1'/><Param nSamplesPerPixel='1'/></Integrator></Scene>")
    (tmp_path / "m.ori").write_text(body)
    hs = oa.HostScene.load_ori(tmp_path / "m.ori")
    m = hs.mats.reshape(-1, oa.MAT_FLOATS)
    f32 = np.float32
    assert m[0][0] == 3 and (m[0][1:4] == f32([0.9, 0.8, 0.7])).all()
    assert m[1][0] == 4 and (m[1][1:4] == f32([0.1, 0.2, 0.3])).all() and (m[1][4:7] == f32(0.4)).all()
    assert m[1][7] == f32(0.12) and m[1][8] == f32(ref.L.ref_roughness_to_alpha(f32(0.12)))
    assert m[2][0] == 5 and (m[2][13:16] == f32([0.6, 0.7, 0.8])).all() and (m[2][10:13] == f32(0.3)).all()
    assert m[2][16] == f32(0.2) and m[2][17] == f32(ref.L.ref_roughness_to_alpha(f32(0.2))) and m[2][18] == f32(1.45)
    assert m[3][0] == 0 and m[3][4] == 20
    assert m[4][0] == 0 and m[4][4] == 0
    # uber needs every texture (the reference dereferences each one)
    (tmp_path / "u.ori").write_text(body.replace(c("Kt", "0.3,0.3,0.3"), ""))
    with pytest.raises(oa.OrionError) as e:
        oa.HostScene.load_ori(tmp_path / "u.ori")
    assert "Kt" in str(e.value)


@pytest.mark.parametrize("body,frag", [
    ("<Scene><Camera type='pinhole'><Param fov='40'/><Param filmsize='4,4'/></Camera>"
     "<Model type='meshHandy'><Param p='0,0,0,1,0,0,0,1,0'/><Param n='0,0,1'/>"
     "<Material type='plastic'/></Model><Integrator type='path'><Param maxDepth='1'/>"
     "<Param nSamplesPerPixel='1'/></Integrator></Scene>", "plastic"),
    ("<Scene><Camera type='pinhole'><Param fov='40'/><Param filmsize='4,4'/></Camera>"
     "<Model type='meshHandy'><Param p='0,0,0,1,0,0,0,1,0'/>"
     "<Material type='matte'><Kd><Param type='constant'/><Param color='1'/></Kd></Material></Model>"
     "<Integrator type='path'><Param maxDepth='1'/><Param nSamplesPerPixel='1'/></Integrator></Scene>",
     "normals"),
    ("<Scene><Camera type='pinhole'>", "XML"),
])
def test_loader_errors(tmp_path, body, frag):
    (tmp_path / "bad.ori").write_text(body)
    with pytest.raises(oa.OrionError) as e:
        oa.HostScene.load_ori(tmp_path / "bad.ori")
    assert frag in str(e.value)
    with pytest.raises(oa.OrionError):
        oa.HostScene.load_ori(tmp_path / "missing.ori")


@pytest.mark.parametrize("name,params", [("cornell", None), ("icosphere", [5, 64, 64]),
                                         ("caustics", [4, 64, 64]),
                                         ("batch_job", [300, 64, 64, 0]),
                                         ("batch_job", [2e5, 64, 64, 1])])
def test_bvh_structure(name, params):
    """the 4-wide quantized BVH: decoded child boxes contain their subtrees,
    every triangle sits in one leaf, depth and stack bound consistent"""
    hs = oa.HostScene.generate(name, seed=3, params=params)
    n_nodes, depth, bound = hs.bvh_check()
    assert n_nodes >= 1 and bound == 3 * (depth + 1)
    # a 4-wide tree over n leaves of <= 8 triangles has at least n / 8 / 3 inner nodes
    assert n_nodes >= hs.n_tris / 24


def test_bvh_tiny_scenes():
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0] + [0, 0, 1] * 3], np.float32)
    for n in (1, 2, 5, 9, 17):
        rng = np.random.default_rng(n)
        pn = np.repeat(tri, n, axis=0)
        pn[:, :9] += rng.random((n, 1)).astype(np.float32)
        mats = np.zeros(16, np.float32)
        cam = oa.camera_record([0, 0, 5, 0, 0, 0, 0, 1, 0], 40, 8, 8)
        hs = oa.HostScene.from_records(pn, np.zeros(n, np.int32), -np.ones(n, np.int32), mats,
                                       np.zeros(0, np.float32), cam)
        nn, depth, bound = hs.bvh_check()
        assert nn >= 1


def test_wave_size_bound_rejected_before_any_device_call():
    """Ray slot s of path p is the 32-bit index s * N + p (s < 3): tunings that
    ask for waves above 2^28 paths are refused with ORION_ERR_UNSUPPORTED, and
    negative values with ORION_ERR_ARG, before a scene or context is touched
    (so this runs without a GPU)."""
    import ctypes
    L = oa.lib()
    big = oa.Tuning(wave_paths=(1 << 28) + 1)
    assert L.orion_render_path(None, 0, 1, 5, ctypes.byref(big), None) == -6
    assert b"2^28" in L.orion_last_error()
    assert L.orion_render_batch(None, 0, None, None, None, ctypes.byref(big), None) == -6
    srv = ctypes.c_void_p()
    assert L.orion_server_create(None, 1 << 30, ctypes.byref(big), ctypes.byref(srv)) == -6
    assert L.orion_render_path(None, 0, 1, 5, ctypes.byref(oa.Tuning(spp_per_wave=(1 << 28) + 1)), None) == -6
    assert L.orion_render_path(None, 0, 1, 5, ctypes.byref(oa.Tuning(wave_paths=-1)), None) == -1
    assert L.orion_render_path(None, 0, 1, 5, ctypes.byref(oa.Tuning(trace_plan=3)), None) == -1
    assert b"trace_plan" in L.orion_last_error()
    # the bound itself is accepted by the check (the null scene is then the error)
    ok = oa.Tuning(wave_paths=1 << 28)
    assert L.orion_render_path(None, 0, 1, 5, ctypes.byref(ok), None) == -1
    assert b"null scene" in L.orion_last_error()
