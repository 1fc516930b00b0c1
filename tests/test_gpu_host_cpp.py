"""The C++ host layer (orion_amd/host/orion_gpu.h: Parser -> Integrator::render
-> getFilm()->getAveraged(), as orion's main.cpp) renders an .ori through the
C ABI; its image equals the oracle's for the same file."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import orion_amd as oa

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, 3)[::-1]


def test_cpp_driver_renders_ori(tmp_path, oracle):
    hs = oa.HostScene.generate("cornell", params=[96, 72])
    ori = tmp_path / "cornell.ori"
    hs.write_ori(ori)
    out = tmp_path / "img.pfm"
    r = subprocess.run([str(ROOT / "orion_amd" / "orion_render"), str(ori), str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    img = read_pfm(out)
    back = oa.HostScene.load_ori(ori)
    ref, cnt, _ = oracle.scene(back).render_path(back.info.spp, back.info.max_depth)
    ref = ref / cnt[..., None].astype(np.float32)  # float / (float)count, as getAveraged
    np.testing.assert_array_equal(img.view(np.int32), ref.view(np.int32))


def test_cpp_driver_writes_bmp_and_exr(tmp_path, oracle):
    """The sky scene through .ori (Envmap PNG written beside it) -> BMP
    (saveAsBmp) and EXR; the EXR holds the averaged film, the BMP its bytes."""
    from test_film_io import read_exr
    hs = oa.HostScene.generate("sky", params=[48, 40])
    ori = tmp_path / "sky.ori"
    hs.write_ori(ori)
    for ext in ("exr", "bmp"):
        r = subprocess.run([str(ROOT / "orion_amd" / "orion_render"), str(ori), str(tmp_path / f"img.{ext}")],
                           capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
    _, img = read_exr((tmp_path / "img.exr").read_bytes())
    back = oa.HostScene.load_ori(ori)
    ref, cnt, _ = oracle.scene(back).render_path(back.info.spp, back.info.max_depth)
    ref = oa.film_average(ref, cnt)
    den = np.maximum(np.abs(ref), 1e-6)
    assert ((np.abs(img - ref) <= 1e-3 * den).all(-1)).mean() >= 0.999
    oa.write_image(tmp_path / "mine.bmp", img)
    assert (tmp_path / "mine.bmp").read_bytes() == (tmp_path / "img.bmp").read_bytes()
