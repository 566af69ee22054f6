"""The oracle's path-trace loop (sampleCosineHemisphere, toWorld and Russian roulette,
src/main.cpp:123-141, 181-193) against the closed forms of tests/closed_forms.py, and
planted-mistake builds of the oracle (ORC_MUTANT in oracle/rtx_oracle.cpp) that each of
those checks must catch."""
import numpy as np
import pytest

import closed_forms as CF
from oracle import oracle as O


def oracle_render(v, t, albedo, emission, cam, W, H, spp, mb, rr, seed, env_mode=0, ibl=None):
    o = O.build(v, t)
    p = O.make_params(W, H, spp, mb, rr, seed=seed, env_mode=env_mode, ibl=ibl)
    return O.render(o, cam, p, albedo, emission, nthreads=8)


def test_lobe_quadrature_matches_closed_form():
    """For n = +y the band energies have closed forms: row r of h covers polar angles
    [r pi/h, (r+1) pi/h] with energy sin^2((r+1) pi/h) - sin^2(r pi/h) above the horizon and
    0 below; columns share the energy equally."""
    h, w = 8, 16
    E = CF.lobe_band_energy((0, 1, 0), h, w)
    rows = [np.sin((r + 1) * np.pi / h) ** 2 - np.sin(r * np.pi / h) ** 2 if r < h // 2 else 0.0 for r in range(h)]
    np.testing.assert_allclose(E.sum(1), rows, atol=2e-6)
    np.testing.assert_allclose(E.sum(0), 1.0 / w, atol=2e-6)
    # any normal: the lobe integrates to 1 and its mean direction is (2/3) n
    n = np.array([0.6, 0.8, 0.0])
    assert abs(CF.lobe_band_energy(n, h, w).sum() - 1.0) < 2e-6


def test_furnace_distribution_reproduces_unbiased_mean():
    mean, var, _ = CF.furnace_moments(6)
    np.testing.assert_allclose(mean, [sum(a ** k for k in range(7)) for a in CF.FURNACE_ALBEDO])
    assert (var > 0).all()


def test_closed_forms_hold_on_oracle():
    results = CF.all_checks(oracle_render)
    for r in results:
        print(r)
    bad = [r for r in results if not r.ok]
    assert not bad, bad


# Which checks each planted mistake must fail (by name prefix).
MUTANTS = {
    1: ("uniform hemisphere sampler", ["sky plane n=", "IBL bands"]),
    2: ("normal in the x slot of toWorld", ["sky plane n=", "IBL bands"]),
    3: ("roulette without throughput /= p", ["roulette furnace mean"]),
    4: ("roulette probability = channel mean", ["roulette furnace blue variance"]),
    5: ("tangent not normalized", ["IBL bands"]),
}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_closed_forms_catch_planted_mistake(k):
    what, must_fail = MUTANTS[k]
    with O.using(O.mutant_path(k)):
        results = CF.all_checks(oracle_render)
    failed = [r for r in results if not r.ok]
    print(f"mutant {k} ({what}): {len(failed)} of {len(results)} checks fail")
    for r in failed:
        print("  ", r)
    for prefix in must_fail:
        assert any(r.name.startswith(prefix) for r in failed), f"mutant {k} ({what}) passes every '{prefix}' check"
