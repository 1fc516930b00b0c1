"""Infinite lights: the PNG reader behind <Envmap name=...>, Envmap and
DistantLight sampling, their power and the world bounding sphere, checked
against the reference's own sources (oracle/_ref: light/envmap.cpp,
light/distant.cpp, texture/imagetexture.cpp + stb_image_aug.c) and against
closed forms."""
import struct
import zlib

import numpy as np
import pytest

import orion_amd as oa

F32 = np.float32


# ------------------------------------------------------------------ PNG
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def write_png(path, px, ctype, palette=None, trns=None):
    """8-bit PNG with every filter type in turn (row y uses filter y % 5)."""
    h, w = px.shape[:2]
    spp = px.shape[2] if px.ndim == 3 else 1
    img = px.reshape(h, w * spp).astype(np.int32)
    raw = bytearray()
    for y in range(h):
        ft = y % 5
        raw.append(ft)
        for x in range(w * spp):
            a = img[y, x - spp] if x >= spp else 0
            b = img[y - 1, x] if y else 0
            c = img[y - 1, x - spp] if (y and x >= spp) else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            raw.append((int(img[y, x]) - pred) & 255)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    out += chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")
    open(path, "wb").write(out)


def ori_with_envmap(d, name=None, xform="", world_up="0,0,1", L="1", tris=True):
    env = (f"<Envmap>{xform}<Param L='{L}'/>" + (f"<Param name='{name}'/>" if name else "") +
           f"<Param nLightSamples='1'/><Param worldUp='{world_up}'/></Envmap>")
    model = ("<Model type='meshHandy'><Param p='-2,0,-1, 3,0,-1, 0,2.5,1.5'/><Param n='0,0,1'/>"
             "<Material type='matte'><Kd><Param type='constant'/><Param color='0.5'/></Kd></Material>"
             "</Model>") if tris else ""
    body = ("<Scene><Resource filename='./'/><Camera type='pinhole'><Transform><Param lookat='0,0,5, 0,0,0, 0,1,0'/>"
            "</Transform><Param fov='40'/><Param filmsize='8,8'/></Camera>" + env + model +
            "<Integrator type='path'><Param maxDepth='3'/><Param nSamplesPerPixel='1'/></Integrator></Scene>")
    p = d / "s.ori"
    p.write_text(body)
    return oa.HostScene.load_ori(p)


@pytest.mark.parametrize("kind", ["rgb", "rgba", "palette", "palette_trns"])
def test_png_reader_matches_stb(tmp_path, ref, kind):
    rng = np.random.default_rng(len(kind))
    h, w = 11, 13
    if kind == "rgb":
        write_png(tmp_path / "e.png", rng.integers(0, 256, (h, w, 3)), 2)
    elif kind == "rgba":
        write_png(tmp_path / "e.png", rng.integers(0, 256, (h, w, 4)), 6)
    else:
        pal = rng.integers(0, 256, (7, 3))
        trns = rng.integers(0, 256, 5) if kind == "palette_trns" else None
        write_png(tmp_path / "e.png", rng.integers(0, 7, (h, w)), 3, palette=pal, trns=trns)
    hs = ori_with_envmap(tmp_path, "e.png")
    mine = hs.env_images[0]
    ch, theirs = ref.load_image(tmp_path / "e.png")
    assert ch in (3, 4)
    assert mine.shape == theirs.shape == (h, w, 3)
    np.testing.assert_array_equal(mine.view(np.int32), theirs.view(np.int32))


def test_png_reader_rejects(tmp_path):
    write_png(tmp_path / "g.png", np.zeros((4, 4), np.uint8), 0)   # grey: 1 channel
    with pytest.raises(oa.OrionError) as e:
        ori_with_envmap(tmp_path, "g.png")
    assert "channel" in str(e.value)
    (tmp_path / "bad.png").write_bytes(b"not a png")
    with pytest.raises(oa.OrionError):
        ori_with_envmap(tmp_path, "bad.png")
    with pytest.raises(oa.OrionError):
        ori_with_envmap(tmp_path, "missing.png")


# ------------------------------------------------------------------ lights
def _m4(r9):
    m = np.eye(4, dtype=F32)
    m[:3, :3] = np.asarray(r9, F32).reshape(3, 3)
    return np.ascontiguousarray(m)


def _close(a, b, rtol=1e-5, atol=1e-6):
    return np.allclose(a, b, rtol=rtol, atol=atol, equal_nan=True)


def _dirs(rng, n):
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)


@pytest.fixture
def env_scene(tmp_path, oracle, ref):
    rng = np.random.default_rng(4)
    px = rng.integers(0, 256, (16, 32, 3))
    px[3:5, 20:23] = 255                       # a bright spot
    write_png(tmp_path / "sky.png", px, 2)
    xform = "<Transform><Param rotate='1,0,0,-90'/><Param rotate='0,0,1,25'/></Transform>"
    hs = ori_with_envmap(tmp_path, "sky.png", xform=xform, world_up="0.2,0.9,0.1", L="2,1.5,1")
    osc = oracle.scene(hs)
    rec = hs.lights.reshape(-1, 32)[-1]
    assert rec[0] == 4
    c, r = osc.world_sphere()
    c = np.ascontiguousarray(c, F32)
    e = ref.L.ref_envmap_new(str(tmp_path / "sky.png").encode(), np.array(rec[2:5], F32), _m4(rec[8:17]),
                             _m4(rec[17:26]), c, r)
    yield hs, osc, e, rec
    ref.L.ref_envmap_free(e)


def test_envmap_orientation_matches_reference(ref):
    import ctypes as C
    for pairs, up in [([("rotate", "1,0,0,-90"), ("rotate", "0,0,1,25")], "0.2,0.9,0.1"),
                      ([], "0,1,0"), ([("rotate", "0,1,0,40")], "0,0,1"), ([], "0,0,-1")]:
        xml = "".join(f"<Param {k}='{v}'/>" for k, v in pairs)
        import tempfile, pathlib
        with tempfile.TemporaryDirectory() as d:
            hs = ori_with_envmap(pathlib.Path(d), None, xform=f"<Transform>{xml}</Transform>" if xml else "",
                                 world_up=up)
        rec = hs.lights.reshape(-1, 32)[-1]
        n = len(pairs)
        keys = (C.c_char_p * max(n, 1))(*[k.encode() for k, _ in pairs])
        vals = (C.c_char_p * max(n, 1))(*[v.encode() for _, v in pairs])
        out = np.zeros(32, F32)
        ref.L.ref_envmap_xform(n, keys, vals, up.encode(), out)
        np.testing.assert_array_equal(rec[8:17], out[:16].reshape(4, 4)[:3, :3].reshape(-1))
        np.testing.assert_array_equal(rec[17:26], out[16:].reshape(4, 4)[:3, :3].reshape(-1))


def test_envmap_matches_reference(env_scene, ref):
    hs, osc, e, rec = env_scene
    L = hs.n_lights - 1
    rng = np.random.default_rng(11)
    # power (envmap.cpp:51-54): the map average x pi r^2
    out = np.zeros(3, F32)
    ref.L.ref_envmap_power(e, out)
    np.testing.assert_array_equal(osc.light_power(L), out)
    # Le and pdf_Li over directions: an ulp of acos / atan2 can move a lookup
    # across a texel edge, so a few directions may differ
    n, le_diff, pdf_diff = 3000, 0, 0
    for d in _dirs(rng, n):
        a = osc.env_le(L, d)
        b = np.zeros(3, F32)
        ref.L.ref_envmap_le(e, d, b)
        le_diff += not np.array_equal(a, b)
        pa, pb = osc.env_pdf_li(L, d), ref.L.ref_envmap_pdf_li(e, d)
        pdf_diff += not _close(pa, pb)
    assert le_diff <= 0.003 * n and pdf_diff <= 0.003 * n, (le_diff, pdf_diff)
    # sample_Li: Li and the map pdf come from the Distribution2D (no trig),
    # directions through sin / cos
    p = np.array([0.1, 0.2, 0.3], F32); ng = np.array([0, 0, 1], F32)
    for _ in range(1000):
        u = rng.random(2).astype(F32)
        a = osc.light_sample_li(L, p, ng, u)
        b = np.zeros(14, F32)
        ref.L.ref_envmap_sample_li(e, p, ng, u, b)
        np.testing.assert_array_equal(a[:3], b[:3])
        assert _close(a[3:7], b[3:7], rtol=1e-4, atol=2e-6), (a, b)
        if a[6] > 0:
            assert _close(a[7:], b[7:], rtol=1e-4, atol=1e-4), (a, b)
    # sample_Le (photons)
    for _ in range(1000):
        u1, u2 = rng.random(2).astype(F32), rng.random(2).astype(F32)
        a = osc.light_sample_le(L, u1, u2)
        b = np.zeros(14, F32)
        ref.L.ref_envmap_sample_le(e, u1, u2, b)
        np.testing.assert_array_equal(a[:3], b[:3])
        assert _close(a[3:], b[3:], rtol=1e-4, atol=1e-4), (a, b)


def test_distant_matches_reference(tmp_path, oracle, ref):
    body = ("<Scene><Resource filename='./'/><Camera type='pinhole'><Transform><Param lookat='0,0,5, 0,0,0, 0,1,0'/>"
            "</Transform><Param fov='40'/><Param filmsize='8,8'/></Camera>"
            "<Model type='meshHandy'><Param p='-2,0,-1, 3,0,-1, 0,2.5,1.5'/><Param n='0,0,1'/>"
            "<Material type='matte'><Kd><Param type='constant'/><Param color='0.5'/></Kd></Material></Model>"
            "<Light type='distant'><Param L='3,2,1'/><Param from='1,2,3'/><Param to='-0.3,4,1.7'/></Light>"
            "<Integrator type='path'><Param maxDepth='3'/><Param nSamplesPerPixel='1'/></Integrator></Scene>")
    (tmp_path / "d.ori").write_text(body)
    hs = oa.HostScene.load_ori(tmp_path / "d.ori")
    osc = oracle.scene(hs)
    rec = hs.lights.reshape(-1, 32)[0]
    assert rec[0] == 3
    c, r = osc.world_sphere()
    c = np.ascontiguousarray(c, F32)
    lt = ref.L.ref_distant_new(b"3,2,1", b"1,2,3", b"-0.3,4,1.7", c, r)
    try:
        d = np.zeros(3, F32)
        ref.L.ref_distant_dir(lt, d)
        np.testing.assert_array_equal(rec[5:8], d)
        out = np.zeros(3, F32)
        ref.L.ref_distant_power(lt, out)
        np.testing.assert_array_equal(osc.light_power(0), out)
        rng = np.random.default_rng(2)
        for _ in range(300):
            p = rng.normal(size=3).astype(F32); ng = _dirs(rng, 1)[0]
            u = rng.random(2).astype(F32)
            a = osc.light_sample_li(0, p, ng, u)
            b = np.zeros(14, F32)
            ref.L.ref_distant_sample_li(lt, p, ng, u, b)
            np.testing.assert_array_equal(a, b)
            u1, u2 = rng.random(2).astype(F32), rng.random(2).astype(F32)
            a = osc.light_sample_le(0, u1, u2)
            ref.L.ref_distant_sample_le(lt, u1, u2, b)
            np.testing.assert_array_equal(a[:3], b[:3])
            assert _close(a[3:], b[3:], rtol=1e-5, atol=1e-5), (a, b)
    finally:
        ref.L.ref_distant_free(lt)


def test_world_sphere_closed_form(oracle):
    hs = oa.HostScene.generate("sky", params=[16, 16])
    c, r = oracle.scene(hs).world_sphere()
    p = hs.tri_pn.reshape(-1, 18)[:, :9].reshape(-1, 3).astype(np.float64)
    lo, hi = p.min(0), p.max(0)
    np.testing.assert_allclose(c, (lo + hi) / 2, rtol=1e-7, atol=1e-7)
    assert abs(r - np.linalg.norm(hi - lo) / 2) <= 1e-6 * r


def test_constant_envmap_closed_forms(oracle):
    """A constant map: Le = L everywhere; the 1 x 1 Distribution2D has pdf 1 in
    (u, v), so pdf_Li = 1 / (2 pi^2 sin theta); power = pi r^2 L."""
    hs = oa.HostScene.generate("cornell", params=[8, 8])
    hs.add_envmap([0.5, 1.0, 2.0])
    osc = oracle.scene(hs)
    l = hs.n_lights - 1
    rng = np.random.default_rng(3)
    for d in _dirs(rng, 200):
        np.testing.assert_array_equal(osc.env_le(l, d), F32([0.5, 1.0, 2.0]))
        theta = np.arccos(np.clip(d[2], -1, 1))
        assert _close(osc.env_pdf_li(l, d), 1 / (2 * np.pi ** 2 * np.sin(theta)), rtol=1e-4)
    for _ in range(200):
        u = rng.random(2).astype(F32)
        a = osc.light_sample_li(l, [0.5, 0.5, 0.5], [0, 1, 0], u)
        wi, pdf = a[3:6], a[6]
        np.testing.assert_array_equal(a[:3], F32([0.5, 1.0, 2.0]))
        # theta = pi u1, phi = 2 pi u0 in the map frame (identity here)
        assert _close(wi, [np.sin(np.pi * u[1]) * np.cos(2 * np.pi * u[0]),
                           np.sin(np.pi * u[1]) * np.sin(2 * np.pi * u[0]), np.cos(np.pi * u[1])], atol=1e-5)
        assert _close(pdf * 2 * np.pi ** 2 * np.sin(np.pi * u[1]), 1.0, rtol=1e-4)
    _, r = osc.world_sphere()
    assert _close(osc.light_power(l), np.pi * r * r * F32([0.5, 1.0, 2.0]), rtol=1e-6)
