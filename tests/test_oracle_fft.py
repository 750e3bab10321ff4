"""Pins the oracle's transforms (chd_oracle.c) against brute force and numpy.

The reference computes these with FFTW (transform_pal_3d.cpp:92-93,201,218;
transform_pal_2d.cpp:88-89), which is not available here; FFTW's r2c/c2r
definitions are the textbook unnormalised DFT, so a brute-force DFT and
numpy's independent pocketfft implementation pin them.
"""
import numpy as np
import pytest


def test_fft3d_matches_bruteforce_dft(orc):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(4, 8, 16))
    got = orc.fft3d(x, np.zeros_like(x), -1)
    want = orc.dft3d_naive(x)
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_dft_naive_matches_numpy(orc):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 8))
    assert np.allclose(orc.dft3d_naive(x), np.fft.fftn(x), atol=1e-10)


@pytest.mark.parametrize("shape", [(8, 32, 16), (1, 16, 32)])
def test_r2c_half_spectrum(orc, shape):
    rng = np.random.default_rng(2)
    x = rng.uniform(-3e4, 3e4, size=shape)
    got = orc.r2c3d(x)
    want = np.fft.rfftn(x, axes=(0, 1, 2))
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("shape", [(8, 32, 16), (1, 16, 32)])
def test_c2r_roundtrip_unnormalised(orc, shape):
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.0, 1.0, size=shape)
    back = orc.c2r3d(orc.r2c3d(x), shape)
    assert np.allclose(back, x * np.prod(shape), atol=1e-9)


@pytest.mark.parametrize("shape", [(8, 32, 16), (1, 16, 32)])
def test_c2r_of_filter_shaped_spectrum(orc, shape):
    """Spectra the filter produces: kx = 0 and Nyquist planes zero, other x
    bins arbitrary.  c2r = unnormalised inverse of the Hermitian extension."""
    rng = np.random.default_rng(4)
    Z, Y, X = shape
    h = rng.normal(size=(Z, Y, X // 2 + 1)) + 1j * rng.normal(size=(Z, Y, X // 2 + 1))
    h[:, :, 0] = 0
    h[:, :, X // 2] = 0
    got = orc.c2r3d(h, shape)
    want = np.fft.irfftn(h, s=shape, axes=(0, 1, 2)) * np.prod(shape)
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


@pytest.mark.parametrize("N", [8, 16, 32])
def test_window_is_a_partition_of_unity(orc, N):
    """The raised-cosine window sums to one at half-tile hops, so the
    overlapping inverse tiles need no synthesis window
    (transform_pal_3d.cpp:60-67)."""
    w = np.array([orc.window(n, N) for n in range(N)])
    assert np.allclose(w[: N // 2] + w[N // 2:], 1.0, atol=1e-15)
    assert np.allclose(w, w[::-1], atol=1e-15)
    assert (w > 0).all() and (w < 1).all()
    # its spectrum lives on bins {0, +-1} only
    W = np.fft.fft(w)
    assert np.abs(W[2:N - 1]).max() < 1e-12
