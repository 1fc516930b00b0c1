"""SPPM pins on the CPU: the oracle's grid (bounds, resolution, per-cell
visible-point counts) recomputed independently in numpy from the visible
points, and structural closed forms of the photon pass."""
import numpy as np
import pytest

import orion_amd as oa


def grid_from_vps(vp, hash_size):
    """ToGrid + hash (sppm.cpp:40-56, 149-193) in float32 / uint32 numpy."""
    vp = vp.reshape(-1, 5)
    valid = vp[:, 4] != 0
    p = vp[valid, :3].astype(np.float32)
    r = vp[valid, 3].astype(np.float32)[:, None]
    lo = (p - r).min(0)
    hi = (p + r).max(0)
    max_r = r.max()
    diag = (hi - lo).astype(np.float32)
    max_diag = np.float32(diag.max())
    base = int(np.float32(max_diag / max_r))
    res = np.maximum((np.float32(base) * diag / max_diag).astype(np.int32), 1)

    def to_grid(q):
        o = (q - lo).astype(np.float32)
        ext = (hi - lo).astype(np.float32)
        o = np.where(hi > lo, o / ext, o).astype(np.float32)
        c = (res.astype(np.float32) * o).astype(np.float32).astype(np.int32)  # trunc toward 0
        return np.clip(c, 0, res - 1)
    cmin = to_grid(p - r)
    cmax = to_grid(p + r)
    counts = np.zeros(hash_size, np.int64)
    pairs = 0
    for a, b in zip(cmin, cmax):
        zz, yy, xx = np.meshgrid(np.arange(a[2], b[2] + 1), np.arange(a[1], b[1] + 1),
                                 np.arange(a[0], b[0] + 1), indexing="ij")
        h = ((xx.astype(np.uint32) * np.uint32(73856093)) ^ (yy.astype(np.uint32) * np.uint32(19349663))
             ^ (zz.astype(np.uint32) * np.uint32(83492791))) % np.uint32(hash_size)
        np.add.at(counts, h.ravel(), 1)
        pairs += h.size
    return lo, hi, res, counts, pairs


@pytest.mark.parametrize("hash_size", [0, 1 << 12])
def test_sppm_grid_matches_numpy(oracle, hash_size):
    hs = oa.HostScene.generate("caustics", params=[3, 48, 40])
    osc = oracle.scene(hs)
    rgb, cnt, dbg = osc.render_sppm(2, 20000, 0.05, 10, hash_size=hash_size, debug_iter=1)
    H = hash_size if hash_size > 0 else 48 * 40
    lo, hi, res, counts, pairs = grid_from_vps(dbg["vp"], H)
    np.testing.assert_array_equal(dbg["grid_bounds"][:3], lo)
    np.testing.assert_array_equal(dbg["grid_bounds"][3:], hi)
    np.testing.assert_array_equal(dbg["grid_info"][:3], res)
    assert dbg["grid_info"][3] == pairs
    np.testing.assert_array_equal(dbg["vp_cell_count"], counts)
    assert counts.sum() == pairs > 0
    assert (cnt == 2).all() and np.isfinite(rgb).all()


def test_sppm_no_deposit_at_depth0(oracle):
    """A lone plane lit by a spot light has no indirect light: photons are only
    deposited at depth > 0 (sppm.cpp:228), so every M and per-cell count is 0."""
    hs = oa.HostScene.generate("caustics", params=[0, 32, 32])
    # keep only the plane (the last two triangles) and its material
    n = hs.n_tris
    keep = slice(n - 2, n)
    tri = hs.tri_pn.reshape(-1, 18)[keep]
    mat = hs.tri_mat[keep]
    h2 = oa.HostScene.from_records(tri, mat, hs.tri_light[keep], hs.mats, hs.lights, hs.camera)
    osc = oracle.scene(h2)
    rgb, cnt, dbg = osc.render_sppm(2, 5000, 0.05, 10, debug_iter=0)
    assert dbg["M"].sum() == 0 and dbg["photon_cell_count"].sum() == 0
    assert dbg["vp_cell_count"].sum() > 0
    # the image is the direct term only: Ld / nIter snapshots (sppm.cpp:313-320)
    assert rgb.sum() > 0
