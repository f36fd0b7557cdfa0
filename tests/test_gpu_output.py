"""GPU output stage against the oracle: mcpt_resolve_rgb8 bit for bit (every
tone map, with planted NaN / inf / negative values) and mcpt_auto_exposure
within FP32 rounding (both sum the same FP32 terms in FP64, in different
orders)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _film_values(shape, seed):
    rng = np.random.default_rng(seed)
    v = rng.lognormal(0.0, 1.5, shape).astype(np.float32)
    v[0, 0] = np.nan
    v[0, 1] = np.inf
    v[0, 2] = -2.0
    v[1, :4] = 0.0
    return v


@pytest.mark.parametrize("tonemap", [0, 1, 2])
def test_resolve_rgb8_matches_oracle(mcpt, ctx, oracle, tonemap):
    W, H, spp = 123, 45, 7
    vals = _film_values((H, W, 3), tonemap)
    film = ctx.film(W, H)
    film.write(vals)
    for exposure in (1.0, 0.37):
        got = film.resolve_rgb8(spp, tonemap=tonemap, exposure=exposure)
        want = oracle.resolve_rgb8(vals, spp, tonemap, exposure)
        assert np.array_equal(got, want), (tonemap, exposure)
    film.close()


def test_auto_exposure_matches_oracle(mcpt, ctx, oracle):
    W, H, spp = 640, 360, 5
    vals = _film_values((H, W, 3), 11)
    film = ctx.film(W, H)
    film.write(vals)
    got = film.auto_exposure(spp)
    want = oracle.auto_exposure(vals, spp)
    assert got == pytest.approx(want, rel=1e-6)
    # constant image: 0.18 / (1e-4 + L)
    film.write(np.full((H, W, 3), 0.5 * spp, np.float32))
    assert film.auto_exposure(spp) == pytest.approx(0.18 / (1e-4 + 0.5), rel=2e-6)
    film.close()
