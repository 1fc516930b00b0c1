"""Batched launch, serving loop, hipGraph replay and autotuner on the GPU.

Each job's film must be the same, bit for bit, whether it is rendered alone,
packed with other jobs into shared wavefronts, replayed from a captured
graph, or served by the request loop: every path's arithmetic depends only
on its own job and sample index, and every film pixel accumulates its
samples in sample order whatever the wave split."""
import time

import numpy as np
import pytest

import orion_amd as oa

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.int32)


@pytest.fixture(scope="module")
def scenes(device):
    out = {}
    out["gold"] = oa.DeviceScene(device, oa.HostScene.generate("batch_job", seed=11, params=[5000, 96, 80, 1]))
    out["glass"] = oa.DeviceScene(device, oa.HostScene.generate("batch_job", seed=12, params=[20000, 64, 64, 2]))
    out["lambert"] = oa.DeviceScene(device, oa.HostScene.generate("batch_job", seed=13, params=[1500, 72, 40, 0]))
    out["cornell"] = oa.DeviceScene(device, oa.HostScene.generate("cornell", params=[48, 48]))
    return out


def job_list(sc):
    # (scene, width, height, spp, max_depth): uneven films, spp and depth,
    # one scene used twice at two film sizes
    return [(sc["gold"], 96, 80, 6, 5), (sc["glass"], 64, 64, 3, 8), (sc["cornell"], 48, 48, 5, 5),
            (sc["gold"], 40, 56, 4, 7), (sc["lambert"], 72, 40, 7, 6)]


def test_batch_alone_equals_mix(device, scenes):
    jobs = job_list(scenes)
    alone = [oa.render_batch(device, [j])[0][0] for j in jobs]
    for wp in (0, 1, 30000):  # default waves, one sample per job per wave, a few
        mixed, st = oa.render_batch(device, jobs, wave_paths=wp)
        assert st.paths == sum(W * H * spp for _, W, H, spp, _ in jobs)
        for (rgb, cnt), (argb, acnt), (_, W, H, spp, _) in zip(mixed, alone, jobs):
            assert (cnt == spp).all()
            np.testing.assert_array_equal(acnt, cnt)
            np.testing.assert_array_equal(bits(argb), bits(rgb))


def test_long_segment_waves_equal_short(device, scenes):
    """Waves of 2^21 paths or more use 16384-path queue segments shaded by
    several blocks each, and a multi-job wave deals every job's segments to
    all XCDs; smaller waves use 512-path segments and one block per segment.
    Films must not depend on which."""
    jobs = [(scenes["gold"], 768, 512, 6, 5), (scenes["glass"], 640, 640, 5, 8),
            (scenes["lambert"], 512, 384, 8, 6)]
    short = [oa.render_batch(device, [j], wave_paths=1 << 18)[0][0] for j in jobs]
    for wp in (0, 1 << 22):  # one wave of all three jobs; waves of ~4M paths mixing two
        mixed, st = oa.render_batch(device, jobs, wave_paths=wp)
        assert st.paths == sum(W * H * spp for _, W, H, spp, _ in jobs)
        for (rgb, cnt), (srgb, scnt), (_, W, H, spp, _) in zip(mixed, short, jobs):
            assert (cnt == spp).all()
            np.testing.assert_array_equal(scnt, cnt)
            np.testing.assert_array_equal(bits(srgb), bits(rgb))
    # one job: 300 samples of a 96x80 film in one 2.3M-path wave vs 30 per wave
    sc = scenes["gold"]
    sc.film_clear()
    sc.render_path(0, 300, 5, spp_per_wave=30)
    a = sc.film()[0].copy()
    sc.film_clear()
    sc.render_path(0, 300, 5, spp_per_wave=300)
    np.testing.assert_array_equal(bits(a), bits(sc.film()[0]))


def test_batch_job_equals_render_path(device, scenes):
    """a batched job whose film is its scene's own equals orion_render_path"""
    sc = scenes["cornell"]
    sc.film_clear()
    sc.render_path(0, 5, 5)
    rgb, cnt = sc.film()
    (brgb, bcnt), = oa.render_batch(device, [(sc, 48, 48, 5, 5)])[0]
    np.testing.assert_array_equal(cnt, bcnt)
    np.testing.assert_array_equal(bits(rgb), bits(brgb))


def test_batch_job_matches_oracle(device, scenes, oracle):
    """the batched path against the oracle, to the radiance contract"""
    from test_gpu_parity import image_compare, MEAN_TOL, PIX_FRAC
    hs = oa.HostScene.generate("batch_job", seed=13, params=[1500, 72, 40, 0])
    ref, rcnt, _ = oracle.scene(hs).render_path(7, 6)
    mixed, _ = oa.render_batch(device, job_list(scenes), wave_paths=1)
    rgb, cnt = mixed[4]
    np.testing.assert_array_equal(cnt, rcnt)
    frac, mean_err, _ = image_compare(rgb, ref)
    assert frac >= PIX_FRAC and mean_err <= MEAN_TOL


def test_graph_replay_equals_eager(device, scenes):
    """captured waves replayed several times, with eager work and film reads
    between replays, give the eager films at every step"""
    sc = scenes["glass"]
    eager = []
    sc.film_clear()
    for k in range(0, 8, 2):
        sc.render_path(k, k + 2, 8, spp_per_wave=2)
        eager.append(sc.film()[0].copy())
    sc.film_clear()
    for i, k in enumerate(range(0, 8, 2)):
        sc.render_path(k, k + 2, 8, spp_per_wave=2, use_graph=True)
        rgb, cnt = sc.film()
        np.testing.assert_array_equal(bits(eager[i]), bits(rgb))
        assert (cnt == k + 2).all()
        scenes["lambert"].render_path(0, 1, 6)  # eager work on the same stream between replays
    jobs = job_list(scenes)
    a, _ = oa.render_batch(device, jobs, wave_paths=1)
    for _ in range(3):
        g, _ = oa.render_batch(device, jobs, wave_paths=1, use_graph=True)
        for (x, _), (y, _) in zip(a, g):
            np.testing.assert_array_equal(bits(x), bits(y))


@pytest.mark.parametrize("budget,policy", [("wide", oa.Server.PACK), ("wide", oa.Server.FAIR),
                                           ("wide", oa.Server.ARRIVAL), ("wide", oa.Server.AGED),
                                           ("one_job", oa.Server.PACK)])
def test_server_results_equal_alone(device, scenes, budget, policy):
    jobs = job_list(scenes)
    alone = [oa.render_batch(device, [j])[0][0] for j in jobs]
    film_bytes = [W * H * 16 for _, W, H, _, _ in jobs]
    # "one_job": no two films fit, so jobs run one after another
    nbytes = 100 * max(film_bytes) if budget == "wide" else min(film_bytes)
    srv = oa.Server(device, nbytes, wave_paths=1)
    srv.set_policy(policy)
    ids = [srv.submit(*j) for j in jobs]
    srv.drain()
    st = srv.stats()
    assert st["jobs_done"] == len(jobs) and st["active_jobs"] == 0 and st["queued_jobs"] == 0
    assert st["hbm_in_use"] == 0
    if budget == "one_job":
        assert st["max_jobs_in_wave"] == 1
    for jid, (argb, acnt) in zip(ids, alone):
        rgb, cnt, (t_sub, t_adm, t_done) = srv.poll(jid)
        assert t_sub <= t_adm <= t_done
        np.testing.assert_array_equal(acnt, cnt)
        np.testing.assert_array_equal(bits(argb), bits(rgb))
    srv.close()


def test_server_continuous_batching(device, scenes):
    """a request that arrives while another runs joins its waves"""
    sc = scenes["gold"]
    big = (sc, 512, 512, 64, 5)  # 64 waves of one sample
    small = (scenes["cornell"], 48, 48, 4, 5)
    alone_small = oa.render_batch(device, [small])[0][0]
    # a wave holds one sample of the big film plus the small job
    srv = oa.Server(device, 1 << 30, wave_paths=(1 << 18) + (1 << 14))
    a = srv.submit(*big)
    t0 = time.time()
    while srv.stats()["waves"] < 2 and time.time() - t0 < 30:
        time.sleep(0.001)
    b = srv.submit(*small)
    srv.drain()
    st = srv.stats()
    assert st["mixed_waves"] >= 1 and st["max_jobs_in_wave"] == 2
    rgb, cnt, tb = srv.poll(b)
    _, _, ta = srv.poll(a)
    assert tb[2] <= ta[2]  # the small job finishes while the big one runs
    np.testing.assert_array_equal(bits(alone_small[0]), bits(rgb))
    srv.close()


def test_server_progressive_peek(device, scenes):
    """previews of a running job are whole sample passes: each equals the job
    rendered alone with that many samples, and the last one is the final film"""
    sc = scenes["gold"]
    job = (sc, 96, 80, 12, 5)
    srv = oa.Server(device, 1 << 30, wave_paths=96 * 80)   # one sample pass per wave
    q = srv.submit(sc, 96, 80, 12, 5)
    seen = []
    while True:
        rgb, cnt, k = srv.peek(q)
        assert (cnt == k).all()
        if k and (not seen or seen[-1][2] != k):
            seen.append((rgb, cnt, k))
        if k == 12:
            break
    srv.drain()
    final, fcnt, _ = srv.poll(q)
    assert seen[-1][2] == 12
    np.testing.assert_array_equal(bits(seen[-1][0]), bits(final))
    for rgb, cnt, k in seen[:4]:
        alone = oa.render_batch(device, [(sc, 96, 80, k, 5)])[0][0]
        np.testing.assert_array_equal(bits(alone[0]), bits(rgb))
    srv.close()


def test_autotune(device, scenes):
    jobs = job_list(scenes)
    best = oa.autotune(device, jobs)
    assert best.wave_paths >= 1 << 20 and best.rays_per_wave >= 64 and best.lds_ring in (8, 16, 32)
    a, _ = oa.render_batch(device, jobs)
    t, _ = oa.render_batch(device, jobs, tuning=best)
    for (x, _), (y, _) in zip(a, t):
        np.testing.assert_array_equal(bits(x), bits(y))
    sc = scenes["glass"]
    sc.film_clear()
    sc.render_path(0, 4, 8)
    plain = sc.film()[0].copy()
    sc.film_clear()
    sc.render_path(0, 4, 8, tuning=best)
    np.testing.assert_array_equal(bits(plain), bits(sc.film()[0]))


def test_context_outlives_its_scenes_and_servers():
    """A context refuses to go (ORION_ERR_BUSY) while a scene or a server made
    on it is alive -- they use its stream and wave runner -- and goes once
    they are destroyed; the scene keeps rendering meanwhile."""
    dev = oa.Device(0)
    ds = oa.DeviceScene(dev, oa.HostScene.generate("cornell", params=[32, 32]))
    srv = oa.Server(dev, 1 << 30)
    L = oa.lib()
    assert L.orion_ctx_destroy(dev.h) == -7  # ORION_ERR_BUSY
    srv.close()
    assert L.orion_ctx_destroy(dev.h) == -7
    ds.film_clear()
    st = ds.render_path(0, 2, 3)
    assert st.paths == 2 * 32 * 32
    ds.close()
    dev.close()
    assert not dev.h
