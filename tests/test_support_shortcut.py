"""The neighbor predicate's shortcut (common.h within_support): deciding
sqrt_rn(d2) /_rn h <= 1 from d2 against h^2 (1 -/+ 1e-6) must give exactly
the full expression's answer.  Checked here in IEEE float32 (numpy) on
d2 values within a few hundred ulps of the boundary and on random
distances, for supports over many binades."""
import numpy as np

F = np.float32


def full(d2, h):
    return (np.sqrt(d2) / h) <= F(1)


def shortcut(d2, h):
    hh = h * h
    lo, hi = hh * F(0.999999), hh * F(1.000001)
    return np.where(d2 < lo, True, np.where(d2 > hi, False, full(d2, h)))


def test_shortcut_equals_full_predicate_near_boundary():
    rng = np.random.default_rng(0)
    h = (10.0 ** rng.uniform(-6, 3, 20000)).astype(F)
    hh = (h.astype(np.float64) ** 2)
    k = np.arange(-400, 401)
    d2 = (hh[:, None] * (1.0 + k[None, :] * 2.0 ** -24)).astype(F)
    hb = np.broadcast_to(h[:, None], d2.shape)
    with np.errstate(over="ignore", under="ignore"):
        a, b = full(d2, hb), shortcut(d2, hb)
    assert np.array_equal(a, b)
    # both outcomes occur in the band, so the comparison is not vacuous
    assert a.any() and (~a).any()


def test_shortcut_equals_full_predicate_random():
    rng = np.random.default_rng(1)
    n = 2_000_000
    h = (10.0 ** rng.uniform(-4, 1, n)).astype(F)
    dx = (rng.uniform(-1.2, 1.2, n) * h).astype(F)
    dy = (rng.uniform(-1.2, 1.2, n) * h).astype(F)
    d2 = dx * dx + dy * dy  # float32: each product rounded, then the sum
    assert np.array_equal(full(d2, h), shortcut(d2, h))
