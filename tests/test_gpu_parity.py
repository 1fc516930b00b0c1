"""GPU parity: the HIP path against the oracle on identical inputs.

Contract: nearest-hit triangle ids (and t, u, v) bit-exact; per-pixel FP32
radiance within 1e-3 relative on >= 99.9% of pixels and mean image radiance
within 1e-4 relative (the tolerances are asserted below)."""
import numpy as np
import pytest

import orion_amd as oa

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3       # per pixel, relative
PIX_FRAC = 0.999     # fraction of pixels that must meet REL_TOL
MEAN_TOL = 1e-4      # relative error of the mean image radiance


def random_rays(hs, n, seed, target_scene=True):
    rng = np.random.default_rng(seed)
    p = hs.tri_pn.reshape(-1, 18)
    lo = p[:, :9].reshape(-1, 3).min(0)
    hi = p[:, :9].reshape(-1, 3).max(0)
    ctr = (lo + hi) / 2
    ext = (hi - lo).max()
    o = ctr + (rng.random((n, 3)) - 0.5) * ext * 1.6
    tgt = ctr + (rng.random((n, 3)) - 0.5) * ext * 0.8
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = o.T
    rays[3:6] = d.T
    rays[6] = 0.0
    rays[7] = np.finfo(np.float32).max
    return rays


def image_compare(gpu, ref):
    """returns (fraction of pixels within REL_TOL, rel. error of the mean, bit-equal fraction)"""
    den = np.maximum(np.abs(ref), 1e-6)
    ok = (np.abs(gpu - ref) <= REL_TOL * den).all(axis=-1)
    mean_err = abs(gpu.mean() - ref.mean()) / max(abs(ref.mean()), 1e-12)
    bits = (gpu.view(np.int32) == ref.view(np.int32)).all(axis=-1)
    return ok.mean(), mean_err, bits.mean()


def test_reciprocal_equals_ieee_division(device):
    """The triangle and slab tests' reciprocal (hardware estimate + one
    Newton-Raphson step) equals IEEE 1 / x for every float with magnitude in
    [2^-125, 2^125): the hit t, u, v and the box tests stay those of a true
    division."""
    assert device.selfcheck_rcp() == 0


@pytest.mark.parametrize("scene,params", [("cornell", None), ("icosphere", [5, 64, 64])])
def test_intersect_closest_bitexact(device, oracle, scene, params):
    hs = oa.HostScene.generate(scene, seed=1, params=params)
    ds = oa.DeviceScene(device, hs)
    osc = oracle.scene(hs)
    rays = random_rays(hs, 20000, seed=7)
    t, prim, u, v, tie = ds.intersect_closest(rays)
    ot, oprim, ou, ov, otie = osc.intersect(np.ascontiguousarray(rays.T), mode=0, brute=True)
    assert (prim >= 0).mean() > 0.3
    np.testing.assert_array_equal(prim, oprim)
    hit = prim >= 0
    np.testing.assert_array_equal(t[hit].view(np.int32), ot[hit].view(np.int32))
    np.testing.assert_array_equal(u[hit].view(np.int32), ou[hit].view(np.int32))
    np.testing.assert_array_equal(v[hit].view(np.int32), ov[hit].view(np.int32))
    np.testing.assert_array_equal(tie, otie)
    occ = ds.intersect_any(rays)
    np.testing.assert_array_equal(occ, (oprim >= 0).astype(np.int32))


def test_intersect_edge_cases(device, oracle):
    """Rays along shared edges / through vertices (ties) and degenerate inputs."""
    hs = oa.HostScene.generate("cornell")
    ds = oa.DeviceScene(device, hs)
    osc = oracle.scene(hs)
    p = hs.tri_pn.reshape(-1, 18)
    # aim at every triangle's vertices and edge midpoints from the camera
    o = hs.camera[:3]
    tg = np.concatenate([p[:, 0:3], p[:, 3:6], p[:, 6:9], (p[:, 0:3] + p[:, 3:6]) / 2,
                         (p[:, 3:6] + p[:, 6:9]) / 2])
    d = tg - o
    n = len(d)
    rays = np.zeros((8, n + 2), np.float32)
    rays[0:3, :n] = o[:, None]
    rays[3:6, :n] = d.T
    rays[7] = np.finfo(np.float32).max
    rays[3:6, n] = [0, 0, 0]          # zero direction: no hit
    rays[0:3, n + 1] = [0.5, 0.5, 0.5]
    rays[3:6, n + 1] = [0, 1, 0]
    rays[7, n + 1] = 0.1              # tmax before the ceiling: no hit
    t, prim, u, v, tie = ds.intersect_closest(rays)
    ot, oprim, ou, ov, otie = osc.intersect(np.ascontiguousarray(rays.T), mode=0, brute=True)
    np.testing.assert_array_equal(prim, oprim)
    np.testing.assert_array_equal(tie, otie)
    assert prim[n] == -1 and prim[n + 1] == -1
    print(f"edge-case rays: {n}, ties counted: {int(tie.sum())}")


def test_path_cornell_full(device, oracle):
    """C1 exactly: 256^2, 16 spp, depth 5, against the oracle."""
    hs = oa.HostScene.generate("cornell")
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    st = ds.render_path(0, 16, 5)
    rgb, cnt = ds.film()
    assert (cnt == 16).all()
    assert st.paths == 256 * 256 * 16
    ref, rcnt, _ = oracle.scene(hs).render_path(16, 5)
    gimg = rgb / cnt[..., None]
    rimg = ref / rcnt[..., None]
    frac, mean_err, bits = image_compare(gimg, rimg)
    print(f"cornell: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL


def test_path_icosphere_crop(device, oracle):
    """C2's 1.31M-triangle mesh at 96^2, 2 spp, depth 8 (RR after bounce 3)."""
    hs = oa.HostScene.generate("icosphere", seed=3, params=[8, 96, 96])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 2, 8)
    rgb, cnt = ds.film()
    ref, rcnt, _ = oracle.scene(hs).render_path(2, 8)
    frac, mean_err, bits = image_compare(rgb / cnt[..., None], ref / rcnt[..., None])
    print(f"icosphere: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL


def test_path_all_materials(device, oracle):
    """Every material kind: Oren-Nayar and Lambert walls, plastic and mirror
    boxes, an uber sphere with all five lobes, rough glass and gold spheres."""
    hs = oa.HostScene.generate("materials")
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 16, 6)
    rgb, cnt = ds.film()
    assert (cnt == 16).all()
    ref, rcnt, _ = oracle.scene(hs).render_path(16, 6)
    frac, mean_err, bits = image_compare(rgb / cnt[..., None], ref / rcnt[..., None])
    print(f"materials: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL


def test_path_sky(device, oracle):
    """Envmap (image, rotated, worldUp +y) + distant light: escaped camera and
    specular rays see the map, MIS over the map, the sun as a delta light."""
    hs = oa.HostScene.generate("sky")
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 16, 6)
    rgb, cnt = ds.film()
    ref, rcnt, _ = oracle.scene(hs).render_path(16, 6)
    frac, mean_err, bits = image_compare(rgb / cnt[..., None], ref / rcnt[..., None])
    print(f"sky: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL


def test_path_constant_envmap(device, oracle):
    """The cornell box opened to a constant envmap (uniform map pdf)."""
    hs = oa.HostScene.generate("cornell", params=[64, 64])
    hs.add_envmap([0.3, 0.4, 0.5])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 8, 5)
    rgb, cnt = ds.film()
    ref, rcnt, _ = oracle.scene(hs).render_path(8, 5)
    frac, mean_err, bits = image_compare(rgb / cnt[..., None], ref / rcnt[..., None])
    print(f"cornell+env: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL


def test_film_averaged_on_device(device):
    """getAveraged on the device equals the host division of the read film."""
    hs = oa.HostScene.generate("cornell", params=[40, 24])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 3, 5)
    rgb, cnt = ds.film()
    np.testing.assert_array_equal(ds.film_averaged().view(np.int32), oa.film_average(rgb, cnt).view(np.int32))


def test_path_waves_equal_single(device):
    """Splitting the sample range into several wavefronts must not change a bit."""
    hs = oa.HostScene.generate("cornell", params=[64, 64])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 8, 5, spp_per_wave=8)
    a, _ = ds.film()
    ds.film_clear()
    ds.render_path(0, 8, 5, spp_per_wave=3)
    b, _ = ds.film()
    np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))


def test_c2_full_size_waves_equal(device):
    """The 1.31M-triangle scene at its full 1024x1024 film: 4 samples in one
    4M-path wave (16384-path queue segments, several shading blocks per
    segment) and one sample per wave (512-path segments) give the same film,
    bit for bit, and every pixel counts its 4 samples."""
    hs = oa.HostScene.generate("icosphere", seed=1, params=[8, 1024, 1024])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    st = ds.render_path(0, 4, 8, spp_per_wave=4)
    a, ca = ds.film()
    assert st.paths == 4 * 1024 * 1024 and st.stack_overflows == 0
    ds.film_clear()
    ds.render_path(0, 4, 8, spp_per_wave=1)
    b, cb = ds.film()
    assert (ca == 4).all() and (cb == 4).all()
    assert np.isfinite(a).all() and (a >= 0).all()
    np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))


def test_c2_default_wave_equals_smaller_waves(device):
    """C2 exactly (1.31M triangles, 1024x1024, 64 spp, depth 8): one default
    wave of 2^26 paths -- the default configuration's largest traversal
    grids (786,432 one-wave blocks per bounce) and spill area -- and four
    16-spp waves give the same film, bit for bit, with no dropped stack
    entries."""
    hs = oa.HostScene.generate("icosphere", seed=1, params=[8, 1024, 1024])
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    st = ds.render_path(0, 64, 8)
    a, ca = ds.film()
    assert st.paths == 64 * 1024 * 1024 and st.stack_overflows == 0
    ds.film_clear()
    st2 = ds.render_path(0, 64, 8, spp_per_wave=16)
    b, cb = ds.film()
    assert st2.stack_overflows == 0
    assert (ca == 64).all() and (cb == 64).all()
    np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))


def test_trace_plan_modes_equal(device):
    """Traversal grid plans change which wave traces which queue chunk, never
    a result: full grids (trace_plan 1), grids sized from the previous call's
    longest segments (0, the default: the second call on the scene uses the
    first one's plan), and one wave per queue segment taking all of its
    segment's chunks (2) give the same film bit for bit and the same ray
    counts, over 2^20-path waves in 2048-path segments."""
    hs = oa.HostScene.generate("icosphere", seed=3, params=[6, 256, 256])
    ds = oa.DeviceScene(device, hs)
    films, rays = [], []
    for mode in (1, 0, 0, 2):
        ds.film_clear()
        st = ds.render_path(0, 32, 8, spp_per_wave=16, queue_segment=2048, trace_plan=mode)
        assert st.stack_overflows == 0 and st.paths == 32 * 256 * 256
        films.append(ds.film())
        rays.append((st.closest_rays, st.shadow_rays, st.node_visits_closest, st.node_visits_shadow))
    for (f, c), r in zip(films[1:], rays[1:]):
        np.testing.assert_array_equal(f.view(np.int32), films[0][0].view(np.int32))
        np.testing.assert_array_equal(c, films[0][1])
        assert r == rays[0]


def _tie_scene(with_lower_duplicate=True, res=48):
    """Lambert plane under a 2 x 2 two-triangle quad light 0.3 above it
    (lights 0 and 1, one sided, facing down), plus a black copy of each light triangle at exactly
    the same vertices: the copy of triangle A comes before it in primitive
    order (lower id), the copy of B after it.  A BSDF-sampled ray towards A
    therefore hits the copy first in the closest-hit order (t, primitive id)
    and carries no MIS term; one towards B hits B itself and does."""
    from test_oracle_integrators import area, matte, tri_record
    up = np.array([0.0, 1.0, 0.0])
    P = [np.array(p, float) for p in ([-3, 0, -3], [-3, 0, 3], [3, 0, 3], [3, 0, -3])]
    c = np.array([0.0, 0.3, 0.0])
    Q = [c + np.array(q) for q in ([-1.0, 0, -1.0], [1.0, 0, -1.0], [1.0, 0, 1.0], [-1.0, 0, 1.0])]
    ln = np.array([0.0, -1.0, 0.0])
    light_tris = []
    for a, b, d in ((Q[0], Q[1], Q[2]), (Q[0], Q[2], Q[3])):
        if np.dot(np.cross(b - a, d - a), ln) < 0:
            b, d = d, b
        light_tris.append(tri_record(a, b, d, ln))
    plane = [tri_record(P[0], P[1], P[2], up), tri_record(P[0], P[2], P[3], up)]
    if with_lower_duplicate:
        # 0: copy of A, 1-2: plane, 3: A, 4: B, 5: copy of B
        pn = [light_tris[0]] + plane + light_tris + [light_tris[1]]
        mat = [1, 0, 0, 1, 1, 1]
        tl = [-1, -1, -1, 0, 1, -1]
        lights = [area(3, (5.0, 4.0, 3.0)), area(4, (5.0, 4.0, 3.0))]
    else:
        # 0-1: plane, 2: A, 3: B, 4: copy of B
        pn = plane + light_tris + [light_tris[1]]
        mat = [0, 0, 1, 1, 1]
        tl = [-1, -1, 0, 1, -1]
        lights = [area(2, (5.0, 4.0, 3.0)), area(3, (5.0, 4.0, 3.0))]
    mats = np.concatenate([matte(0.6, 0.5, 0.7), matte(0.0)])
    cam = oa.camera_record([0.0, 0.2, 1.2, 0, 0, 0, 0, 1, 0], 60.0, res, res)
    return oa.HostScene.from_records(np.array(pn, np.float32), np.array(mat, np.int32), np.array(tl, np.int32),
                                     mats, np.array(lights, np.float32), cam)


def test_path_mis_light_ties(device, oracle):
    """MIS visibility queries under exact ties: a light triangle and a black
    copy at the same vertices.  The GPU film equals the oracle's (whose MIS
    term needs the closest hit to be the light's own triangle), and removing
    the lower-id copy changes the image, so the tie decides the result."""
    hs = _tie_scene(True)
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    ds.render_path(0, 32, 3)
    rgb, cnt = ds.film()
    ref, rcnt, _ = oracle.scene(hs).render_path(32, 3)
    frac, mean_err, bits = image_compare(rgb / cnt[..., None], ref / rcnt[..., None])
    print(f"mis ties: within 1e-3 {frac:.6f}, mean rel err {mean_err:.2e}, bit-equal {bits:.6f}")
    assert frac >= PIX_FRAC
    assert mean_err <= MEAN_TOL
    hs2 = _tie_scene(False)
    ds2 = oa.DeviceScene(device, hs2)
    ds2.film_clear()
    ds2.render_path(0, 32, 3)
    rgb2, cnt2 = ds2.film()
    gap = abs((rgb2 / cnt2[..., None]).mean() - (rgb / cnt[..., None]).mean()) / (rgb2 / cnt2[..., None]).mean()
    print(f"mis ties: image mean without the lower-id copy differs by {gap:.3%}")
    assert gap > 0.01
    ref2, rcnt2, _ = oracle.scene(hs2).render_path(32, 3)
    frac2, mean_err2, _ = image_compare(rgb2 / cnt2[..., None], ref2 / rcnt2[..., None])
    assert frac2 >= PIX_FRAC
    assert mean_err2 <= MEAN_TOL
