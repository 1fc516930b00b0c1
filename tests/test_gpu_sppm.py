"""SPPM on the GPU against the oracle: visible-point cell keys (per-cell
counts), per-cell photon counts, per-pixel photon counts M and the grid are
bit-exact; the FP32 image within 1e-3 relative on >= 99.9% of pixels and the
mean within 1e-4 (photon flux Phi is summed with float atomics, so its order
differs between runs; every other quantity is deterministic)."""
import numpy as np
import pytest

import orion_amd as oa

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3
PIX_FRAC = 0.999
MEAN_TOL = 1e-4


def _compare(device, oracle, params, iters, photons, radius, hash_size, debug_iter, scene="caustics",
             depth=10):
    hs = oa.HostScene.generate(scene, params=params)
    ds = oa.DeviceScene(device, hs)
    ds.film_clear()
    st, gd = ds.render_sppm(iters, photons, radius, depth, hash_size=hash_size, debug_iter=debug_iter)
    g_rgb, g_cnt = ds.film()
    o_rgb, o_cnt, od = oracle.scene(hs).render_sppm(iters, photons, radius, depth, hash_size=hash_size,
                                                    debug_iter=debug_iter)
    np.testing.assert_array_equal(gd["grid_info"][:4], od["grid_info"][:4])
    np.testing.assert_array_equal(gd["grid_bounds"], od["grid_bounds"])
    np.testing.assert_array_equal(gd["vp_cell_count"], od["vp_cell_count"])
    np.testing.assert_array_equal(gd["photon_cell_count"], od["photon_cell_count"])
    np.testing.assert_array_equal(gd["M"], od["M"])
    assert (g_cnt == iters).all() and (o_cnt == iters).all()
    g = g_rgb / g_cnt[..., None]
    o = o_rgb / o_cnt[..., None]
    ok = (np.abs(g - o) <= REL_TOL * np.maximum(np.abs(o), 1e-6)).all(-1).mean()
    mean_err = abs(g.mean() - o.mean()) / max(abs(o.mean()), 1e-12)
    print(f"sppm {scene} {params} iters={iters} photons={photons} hash={hash_size}: pairs={gd['grid_info'][3]} "
          f"M={gd['M'].sum()} photon-cells={gd['photon_cell_count'].sum()} within1e-3={ok:.6f} "
          f"mean_err={mean_err:.2e} deposits={st.deposits}")
    assert gd["M"].sum() > 0 and gd["photon_cell_count"].sum() > 0
    assert ok >= PIX_FRAC
    assert mean_err <= MEAN_TOL
    return st


@pytest.mark.parametrize("hash_size", [0, 1 << 22])
def test_sppm_reduced(device, oracle, hash_size):
    _compare(device, oracle, [5, 96, 96], 4, 1 << 16, 0.02, hash_size, 3)


def test_sppm_c3_single_iteration(device, oracle):
    """C3's sizes for one iteration: 512^2, 2^20 photons, r0 0.01, 2^22 cells."""
    _compare(device, oracle, [6, 512, 512], 1, 1 << 20, 0.01, 1 << 22, 0)


def test_sppm_all_materials(device, oracle):
    """Every material kind (Oren-Nayar, plastic, mirror, uber, rough glass,
    gold) under photon transport mode, area + point lights."""
    _compare(device, oracle, [96, 96], 4, 1 << 16, 0.02, 0, 3, scene="materials", depth=6)


def test_sppm_sky(device, oracle):
    """Photons from an envmap and a distant light (world-disk emission), the
    camera pass adding the map on escape at any depth."""
    _compare(device, oracle, [64, 64], 3, 1 << 16, 0.05, 0, 2, scene="sky", depth=6)


def test_sppm_c3_three_iterations(device, oracle):
    """C3's sizes over three iterations: iteration 2's grid comes from radii
    the first two updates shrank, and the camera passes of iterations 1 and 2
    ran on the second stream beside the previous iteration's photon pass
    (sppm.cpp:79-321)."""
    _compare(device, oracle, [6, 512, 512], 3, 1 << 20, 0.01, 1 << 22, 2)


def test_sppm_batches_reuse_buffers(device, oracle):
    """Twenty iterations at 128^2: the camera passes run in batches of 8
    iterations (16 visible-point sets and Ld snapshots, so iteration 17
    reuses iteration 1's set) and the photons in batches of 16 iterations
    (two hit buffers); iteration 17's grid, deposits and M equal the
    oracle's bit for bit and the film matches."""
    _compare(device, oracle, [6, 128, 128], 20, 1 << 14, 0.02, 0, 17)


def test_sppm_c3_sorted_pairs_and_cell_ranges(device, oracle):
    """The radix-sorted (hash key, pixel) pairs and every cell's [start, end)
    equal the reference's per-cell lists (sppm.cpp:170-193) concatenated by
    key, bit for bit, hash collisions' repeated pixels included (C3 sizes,
    iteration 1, after one radius update)."""
    hs = oa.HostScene.generate("caustics", params=[6, 512, 512])
    ds = oa.DeviceScene(device, hs)
    _, gd = ds.render_sppm(2, 1 << 20, 0.01, 10, hash_size=1 << 22, debug_iter=1, pairs=True)
    _, _, od = oracle.scene(hs).render_sppm(2, 1 << 20, 0.01, 10, hash_size=1 << 22, debug_iter=1, pairs=True)
    gp, op = gd["pairs"], od["pairs"]
    print(f"pairs: {len(gp)} (oracle {len(op)}), distinct keys {len(np.unique(gp[:, 0]))}")
    assert len(gp) > 100000
    np.testing.assert_array_equal(gp, op)
    counts = od["vp_cell_count"].astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    want = np.stack([np.where(counts > 0, starts, 0), np.where(counts > 0, starts + counts, 0)], 1)
    np.testing.assert_array_equal(gd["cell_range"], want)


def test_sppm_ragged_photon_segments(device, oracle):
    """Photon counts that are multiples of neither the photon queue segment
    (16,384 entries) nor a block: 3 iterations of 50,000 photons (one batch
    of 150,000, its last segment partly filled) and 4 of 3,001 (a batch
    smaller than one segment, its last block partly filled)."""
    _compare(device, oracle, [5, 96, 96], 3, 50000, 0.02, 0, 2)
    _compare(device, oracle, [5, 96, 96], 4, 3001, 0.02, 0, 3)


def test_sppm_planned_traversal_grids(device, oracle):
    """Twenty-four iterations of 2^19 photons at 96^2: three photon batches
    and three camera batches of 8 iterations each.  The later photon batches
    trace each bounce on a grid sized from the previous batch's photon
    counts, and the third camera batch on grids sized from the first one's
    queue lengths (waves stride over their queue's chunks, so any grid traces
    every ray); iteration 20's grid, deposits and M equal the oracle's bit
    for bit."""
    st = _compare(device, oracle, [5, 96, 96], 24, 1 << 19, 0.02, 0, 20)
    assert st.photon_batch == 8
