"""The device draw scheme against the compiled reference, on the CPU.

The oracle's Philox streams and direct azimuth (rng=1, the scheme the GPU
kernels use) are compared with the reference's std::mt19937 histories
(tests/golden/stat_*.npz) by the two-sample criteria of tally_stats: the
scalars (vacancies per ion, collisions per history, Eel/E0, Enuc/E0, mean
range) within 2 sigma, the range and lateral histograms by a chi-square
homogeneity test, the vacancy and replacement depth histograms by
Hotelling's T^2.  The GPU runs the same comparison at larger samples
(tests/test_gpu_stats.py).  A planted error in the oracle (the azimuth drawn
over half the circle) must fail the same comparison."""
import pytest

import oracle_lib as O
import tally_stats as S
from fixture_cases import stat_cases

# histories of the oracle sample per case (32 equal batches), about 1e7-6e7
# collisions each: heavy-tailed cascade sizes (2 MeV C -> W, the buried wire)
# need thousands of histories before the batch variances settle
N_ORACLE = {"c1": 4000, "c2": 960, "c3": 640, "c4mix": 1920, "c5": 6400, "c2_primaries_kp": 3200,
            "vac_cu_10keV": 3200, "wire": 960, "bwire": 960}
K = 32


def oracle_batches(name, seed=77):
    prob = stat_cases()[name]
    prob = prob.with_(seed=seed).batch(0, N_ORACLE[name]) if prob.primaries.uniform else \
        prob.batch(0, N_ORACLE[name])
    return S.run_batches(lambda b: O.run(b, rng=1, nthreads=8), prob, K)


@pytest.mark.parametrize("name", sorted(N_ORACLE))
def test_device_scheme_matches_reference_tallies(name):
    ref = S.load_ref(name)
    fails, lines = S.check(name, ref, oracle_batches(name))
    print("\n".join(lines))
    assert not fails, fails


def test_planted_azimuth_error_is_detected():
    """An azimuth drawn over [0, pi) instead of [0, 2 pi) shifts the lateral
    spread of every flight; the lateral histograms must catch it."""
    ref = S.load_ref("c1")
    O.plant_fault(1)
    try:
        oth = oracle_batches("c1")
    finally:
        O.plant_fault(0)
    fails, lines = S.check("c1", ref, oth)
    print("\n".join(lines))
    assert "lat_y_hist" in fails or "lat_z_hist" in fails, fails
