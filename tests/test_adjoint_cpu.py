"""Host-side parts of the gradient workflow (CPU only)."""
import numpy as np

from lbmcl_amd.adjoint import gaussian_random_field


def test_gaussian_random_field_is_seeded_and_scaled():
    a = gaussian_random_field(32, 8.0, 1e-6, seed=4)
    b = gaussian_random_field(32, 8.0, 1e-6, seed=4)
    c = gaussian_random_field(32, 8.0, 1e-6, seed=5)
    assert a.shape == (3, 32 ** 3) and np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.allclose(a.std(axis=1), 1e-6, rtol=1e-12)


def test_gaussian_random_field_correlation_length():
    """The filter exp(-k^2 l^2 / 2) gives a Gaussian autocorrelation exp(-r^2 / (4 l^2))."""
    n, ell = 64, 4.0
    f = gaussian_random_field(n, ell, 1.0, seed=9, components=1)[0].reshape(n, n, n)
    for r in (2, 4, 8):
        rho = np.mean(f * np.roll(f, r, axis=2)) / np.mean(f * f)
        assert abs(rho - np.exp(-r * r / (4 * ell * ell))) < 0.08, (r, rho)
