"""Pins the oracle's FFT / Freq / ConjugateNode (pfutil/fftWrap.go:16-95).

Checks that a plausible mistake would fail: a wrong twiddle sign or index, a
dropped normalisation, a transposed axis order.
"""
import itertools

import numpy as np
import pytest


def brute_dft(x, dims, sign):
    """Textbook O(N^2) multi-dimensional DFT over FFTW's row-major layout."""
    x = np.asarray(x).reshape(dims)
    out = np.zeros_like(x, dtype=np.complex128)
    grids = [np.arange(d) for d in dims]
    for k in itertools.product(*grids):
        acc = 0.0 + 0.0j
        for n in itertools.product(*grids):
            ph = sum(kk * nn / d for kk, nn, d in zip(k, n, dims))
            acc += x[n] * np.exp(sign * 2j * np.pi * ph)
        out[k] = acc
    return out.reshape(-1)


@pytest.mark.parametrize("dims", [[4, 4], [3, 5], [8, 2], [2, 3, 4], [6, 6]])
def test_fft_matches_brute_force(oracle_lib, dims):
    rng = np.random.default_rng(sum(dims))
    x = rng.standard_normal(np.prod(dims)) + 1j * rng.standard_normal(np.prod(dims))
    for inverse, sign in ((False, -1), (True, +1)):
        got = oracle_lib.fft(x, dims, inverse=inverse)
        ref = brute_dft(x, dims, sign)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dims", [[8, 16], [16, 8], [9, 9], [12, 20], [7, 11], [8, 8, 8], [6, 10, 12], [32, 32, 32]])
def test_fft_matches_numpy(oracle_lib, dims):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(np.prod(dims)) + 1j * rng.standard_normal(np.prod(dims))
    got = oracle_lib.fft(x, dims)
    ref = np.fft.fftn(x.reshape(dims)).reshape(-1)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * np.sqrt(np.prod(dims))
    goti = oracle_lib.fft(x, dims, inverse=True)
    refi = np.fft.ifftn(x.reshape(dims)).reshape(-1) * np.prod(dims)
    assert np.abs(goti - refi).max() <= 1e-13 * np.abs(refi).max() * np.sqrt(np.prod(dims))


def test_fft_known_answers(oracle_lib):
    dims = [8, 16]
    n = 128
    delta = np.zeros(n, complex)
    delta[0] = 1.0
    assert np.allclose(oracle_lib.fft(delta, dims), np.ones(n), atol=1e-15)
    # single mode exp(+2 pi i (2 r / 8 + 3 c / 16)) -> N * delta at (2, 3)
    r, c = np.divmod(np.arange(n), 16)
    mode = np.exp(2j * np.pi * (2 * r / 8 + 3 * c / 16))
    out = oracle_lib.fft(mode, dims)
    expect = np.zeros(n, complex)
    expect[2 * 16 + 3] = n
    assert np.abs(out - expect).max() < 1e-12
    # round trip divides by N exactly as Euler.Step does (euler.go:42-45)
    x = np.random.default_rng(1).standard_normal(n).astype(complex)
    back = oracle_lib.fft(oracle_lib.fft(x, dims), dims, inverse=True) / n
    assert np.abs(back - x).max() < 1e-14


def test_fftw_wrapper_consistency_ramp(oracle_lib):
    """TestFFTWWrapConsistency (fftwWrap_test.go:11-57): ramp on 8x16."""
    dims = [8, 16]
    data = np.arange(128).astype(complex)
    got = oracle_lib.fft(data, dims)
    ref = np.fft.fft2(data.reshape(dims)).reshape(-1)
    assert np.abs(got - ref).max() < 1e-6
    assert np.abs(oracle_lib.fft(got, dims, inverse=True) - ref.size * np.fft.ifft2(ref.reshape(dims)).reshape(-1)).max() < 1e-6


@pytest.mark.parametrize("dims", [[8, 8], [9, 9], [8, 8, 8], [9, 9, 9]])
def test_conjugate_node(oracle_lib, dims):
    """TestConjugateNode (fftwWrap_test.go:59-90)."""
    for j in range(int(np.prod(dims))):
        f1 = oracle_lib.freq(dims, j)
        f2 = oracle_lib.freq(dims, oracle_lib.conjugate_node(dims, j))
        for a, b in zip(f1, f2):
            assert not (abs(a + b) > 1e-10 and abs(a) < 0.5 - 1e-10)


def test_freq_closed_form(oracle_lib):
    """Freq (fftWrap.go:57-74): components (row, col[, depth]) in (-0.5, 0.5]."""
    dims = [8, 16]
    for i in range(128):
        r, c = divmod(i, 16)
        fr = r / 8 if r / 8 <= 0.5 else r / 8 - 1
        fc = c / 16 if c / 16 <= 0.5 else c / 16 - 1
        assert oracle_lib.freq(dims, i) == pytest.approx([fr, fc], abs=0)
    dims = [4, 4, 4]
    for i in range(64):
        d, rem = divmod(i, 16)
        r, c = divmod(rem, 4)
        w = lambda v: v / 4 if v / 4 <= 0.5 else v / 4 - 1
        assert oracle_lib.freq(dims, i) == pytest.approx([w(r), w(c), w(d)], abs=0)
