"""Each quirk pin of the oracle fails when its quirk is flipped.

The oracle is rebuilt with -DORACLE_MUTATION=k (oracle/oracle.hpp lists the
mutations), loaded beside the shipped library, and the pin that guards quirk k
is run against it: it must raise.  Mutation 0 (the shipped oracle, rebuilt the
same way) must pass every pin, so a failure is the mutation's doing.
"""
import concurrent.futures as cf
import subprocess
from pathlib import Path

import pytest

from tests import oracle_lib as ol
from tests import test_oracle_core as core

ROOT = Path(__file__).resolve().parent.parent

PINS = {
    "pbr_mirror": core.pin_pbr_material_mirror_symmetric,
    "source_lod": core.pin_prefilter_source_lod,
    "hash_rng": core.pin_hash_rng_words,
    "estimator": core.pin_hammersley_vs_reference_estimator,
}
# mutation -> the pin that must catch it (descriptions in oracle/oracle.hpp)
MUTATIONS = {
    1: "pbr_mirror",  # H = normalize(-V + L)
    2: "pbr_mirror",  # direct-light alpha = roughness
    3: "pbr_mirror",  # direct-light Fresnel at NdotV
    4: "source_lod",  # LOD without the factor 0.5
    5: "source_lod",  # saTexel without the 6 faces
    6: "source_lod",  # pdf without + 0.00001
    7: "estimator",  # prefilter alpha = roughness^2
    8: "hash_rng",  # q.y >> 2
    9: "hash_rng",  # seed.yx not swapped
}


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_mutations")

    def build(k):
        so = out / f"liboracle_m{k}.so"
        subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
                        f"-DORACLE_MUTATION={k}", str(ROOT / "oracle" / "oracle_capi.cpp"), "-o", str(so)],
                       check=True, capture_output=True)
        return k, so

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, [0, *MUTATIONS]))


def _run_pin(so, pin):
    saved = ol._L
    ol._L = ol.load(so)
    try:
        pin()
    finally:
        ol._L = saved


def test_unmutated_rebuild_passes_every_pin(variants):
    for pin in PINS.values():
        _run_pin(variants[0], pin)


@pytest.mark.parametrize("k", sorted(MUTATIONS))
def test_pin_catches_mutation(variants, k):
    with pytest.raises(AssertionError):
        _run_pin(variants[k], PINS[MUTATIONS[k]])


def test_source_lod_pin_also_catches_prefilter_alpha(variants):
    # Q4 moves the one-sample LOD as well (D(1) = 1 / (pi a2)).
    with pytest.raises(AssertionError):
        _run_pin(variants[7], core.pin_prefilter_source_lod)
