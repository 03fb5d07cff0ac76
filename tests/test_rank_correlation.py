"""The surrogate's Spearman rank correlation against SciPy's."""
import numpy as np
import pytest
from scipy.stats import spearmanr

from brkga_fft2_amd.surrogate import spearman


@pytest.mark.parametrize("seed", range(6))
def test_spearman_matches_scipy_with_ties(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 400))
    # few distinct values so that ties are common in both inputs
    a = rng.integers(0, max(2, n // 7), n).astype(np.float32)
    b = np.round(rng.normal(size=n) + 0.3 * a, 1)
    if np.all(a == a[0]) or np.all(b == b[0]):
        pytest.skip("constant draw")
    assert abs(spearman(a, b) - spearmanr(a, b).correlation) <= 1e-12


def test_spearman_special_cases():
    x = np.arange(10, dtype=np.float64)
    assert spearman(x, x) == pytest.approx(1.0, abs=1e-15)
    assert spearman(x, -x) == pytest.approx(-1.0, abs=1e-15)
    assert spearman(x, x ** 3) == pytest.approx(1.0, abs=1e-15)   # ranks, not values
    assert np.isnan(spearman(np.ones(5), x[:5]))                  # constant input
    assert np.isnan(spearman(x[:2], x[:2]))                       # fewer than 3 values
