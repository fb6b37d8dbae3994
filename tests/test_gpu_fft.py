"""FourierTransform drop-in (pfutil/fftWrap.go:16-39) on the GPU against the oracle FFT.

Tolerances: FP64 path 1e-13 * sqrt(N) relative to max|X|; FP32 path 2e-6 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [[8, 16], [16, 8], [16, 16], [9, 9], [12, 20], [6, 10], [64, 64], [128, 256], [8, 8, 8],
         [24, 24, 24], [32, 32, 32], [12, 16, 8], [192, 192], [512, 512]]


@pytest.mark.parametrize("dims", SIZES)
def test_fft_fp64_matches_oracle(oracle_lib, dims):
    import gopf_amd as G
    print("fft fp64", dims, flush=True)
    rng = np.random.default_rng(len(dims) * 100 + dims[0])
    n = int(np.prod(dims))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ft = G.FourierTransform(dims)
    for inverse in (False, True):
        ref = oracle_lib.fft(x, dims, inverse=inverse)
        got = x.copy()
        (ft.IFFT if inverse else ft.FFT)(got)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13 * np.sqrt(n), (dims, inverse, err)


@pytest.mark.parametrize("dims", [[16, 16], [64, 64], [32, 32, 32], [256, 256]])
def test_fft_fp32(oracle_lib, dims):
    import gopf_amd as G
    print("fft fp32", dims, flush=True)
    rng = np.random.default_rng(5)
    n = int(np.prod(dims))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ft = G.FourierTransform(dims, precision="fp32")
    got = x.copy()
    ft.FFT(got)
    ref = oracle_lib.fft(x, dims)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-6


def test_fftw_wrap_consistency_ramp(oracle_lib):
    """TestFFTWWrapConsistency (pfutil/fftwWrap_test.go:11-57): 8x16 ramp, FFT then IFFT."""
    import gopf_amd as G
    dims = [8, 16]
    data = np.arange(128).astype(np.complex128)
    ft = G.FourierTransform(dims)
    got = ft.FFT(data.copy())
    ref = oracle_lib.fft(data, dims)
    assert np.abs(got - ref).max() < 1e-6
    back = ft.IFFT(got.copy())
    assert np.abs(back - oracle_lib.fft(ref, dims, inverse=True)).max() < 1e-6
    assert np.abs(back / 128 - data).max() < 1e-12
    for i in range(128):
        assert ft.Freq(i) == oracle_lib.freq(dims, i)


# compiled radix-5 / radix-7 / 2^11 schedules, and lengths without one (Bluestein, up to
# 1024, primes included): FFTW plans any n (pfutil/fftWrap.go:16-23)
ANY_SIZES = [[80, 100], [160, 16], [320, 8], [2048, 4], [9, 15], [97, 16], [1000, 4], [112, 224], [250, 8],
             [1009, 2], [10, 30, 7], [80, 80, 80], [65, 66, 67]]


@pytest.mark.parametrize("dims", ANY_SIZES)
def test_fft_any_length_fp64_matches_oracle(oracle_lib, dims):
    import gopf_amd as G
    print("fft any length fp64", dims, flush=True)
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ft = G.FourierTransform(dims)
    for inverse in (False, True):
        ref = oracle_lib.fft(x, dims, inverse=inverse)
        got = x.copy()
        (ft.IFFT if inverse else ft.FFT)(got)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13 * np.sqrt(n), (dims, inverse, err)
    # and numpy's transform, an independent implementation (sign -1, unnormalised)
    got = x.copy()
    ft.FFT(got)
    ref = np.fft.fftn(x.reshape(dims)).reshape(-1)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13 * np.sqrt(n)


@pytest.mark.parametrize("dims", [[100, 100], [97, 64], [2048, 8], [112, 112]])
def test_fft_any_length_fp32(oracle_lib, dims):
    import gopf_amd as G
    print("fft any length fp32", dims, flush=True)
    rng = np.random.default_rng(7)
    n = int(np.prod(dims))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ft = G.FourierTransform(dims, precision="fp32")
    got = x.copy()
    ft.FFT(got)
    ref = oracle_lib.fft(x, dims)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 5e-6

