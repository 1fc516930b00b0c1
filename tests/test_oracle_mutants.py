"""The oracle's pins can fail: each mutant is the oracle with one planted,
plausible misreading of the reference (oracle/oracle.cpp, ORACLE_MUTANT = k),
built here into a temporary directory, and at least one of the closed-form
or reference checks that pin the oracle must reject it.  CPU only; the
default build (the oracle itself) is untouched."""
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

import test_oracle_integrators as integ
import test_oracle_pins as pins
from oracle_lib import Oracle, have_ref

ROOT = Path(__file__).resolve().parents[1]

# k -> (the planted mistake, the checks expected to catch it)
MUTANTS = {
    1: ("cosine at the shading point dropped from the light-sampled term",
        [lambda o: integ.test_direct_light_on_plane_matches_irradiance_integral(o)]),
    2: ("BSDF-sampled MIS term counted without its weight",
        [lambda o: integ.test_path_furnace_lambert(o, 0.5, 5, 256)]),
    3: ("light-selection pdf ignored",
        [lambda o: integ.test_path_furnace_lambert(o, 0.5, 5, 256)]),
    4: ("path depth off by one",
        [lambda o: integ.test_path_furnace_mirror_deterministic(o, 0.5, 3)]),
    5: ("triangle area pdf halved",
        [lambda o: integ.test_area_light_sample_closed_forms(o)]),
    6: ("triangle barycentrics b0 and b1 swapped",
        [lambda o: integ.test_area_light_sample_closed_forms(o)]),
    7: ("emission after a specular bounce dropped",
        [lambda o: integ.test_path_furnace_mirror_deterministic(o, 0.5, 3)]),
    8: ("photon power without the light-selection pdf",
        [lambda o: integ.test_sppm_furnace(o, 0.5, 4)]),
    9: ("SPPM radius reduction with alpha 3/4",
        [lambda o: integ.test_sppm_radius_update(o)]),
    10: ("SPPM film as the running average instead of orion's snapshot mean",
         [lambda o: integ.test_sppm_direct_light_on_plane(o)]),
    11: ("photons deposit at their first hit",
         [lambda o: integ.test_sppm_direct_light_on_plane(o)]),
    12: ("offsetRayOrigin without its one-ulp nudge",
         [lambda o: pins.test_spawn_matches_reference(o, _ref())]),
    13: ("Distribution1D::sampleDiscrete pdf without the 1 / n",
         [lambda o: pins.test_distribution1d_matches_reference(o, _ref()),
          lambda o: integ.test_path_furnace_lambert(o, 0.5, 5, 256)]),
}
NEEDS_REF = {12}


def _ref():
    from oracle_lib import Ref
    return Ref()


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_mutants")
    flags = ["-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-w"]

    def build(k):
        so = out / f"liboracle_mutant{k}.so"
        subprocess.run(["g++", *flags, f"-DORACLE_MUTANT={k}", str(ROOT / "oracle" / "oracle.cpp"), "-o", str(so)],
                       check=True, capture_output=True)
        return k, so
    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, sorted(MUTANTS)))


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_is_caught(mutant_libs, k):
    if k in NEEDS_REF and not have_ref():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    what, checks = MUTANTS[k]
    o = Oracle(mutant_libs[k])
    caught = []
    for chk in checks:
        try:
            chk(o)
            caught.append(False)
        except AssertionError:
            caught.append(True)
        except FileNotFoundError:  # a check against oracle/_ref, which is not built
            continue
    assert any(caught), f"mutant {k} ({what}) passes every check meant to catch it"
