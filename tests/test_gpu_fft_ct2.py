"""GPU drop-ins for ct2's own FFT API (fft.hpp:18-29) in ct2's conventions --
bit-reversed 1-D input, half-complex real output, 4-lane interleaving, the
planar 3-D half spectrum -- against the FP64 oracle (itself pinned by numpy
golden vectors and the DFT's invariants)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N", [2, 8, 32, 256, 1024, 4096])
@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("dir", [1, -1])
def test_fft_complex_1d(ct2, orc, N, lanes, dir):
    batch = 3
    X = np.random.default_rng(N + lanes).standard_normal(batch * 2 * lanes * N)
    got = ct2.fft_complex(X, N, dir, lanes)
    want = np.concatenate([orc.fft_complex(g, N, dir, lanes) for g in np.split(X, batch)])
    assert relerr(got, want) < 1e-14


@pytest.mark.parametrize("N", [2, 8, 64, 512, 4096])
@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("dir", [1, -1])
def test_fft_real_1d(ct2, orc, N, lanes, dir):
    batch = 5
    X = np.random.default_rng(7 * N + lanes).standard_normal(batch * lanes * N)
    got = ct2.fft_real(X, N, dir, lanes)
    want = np.concatenate([orc.fft_real(g, N, dir, lanes) for g in np.split(X, batch)])
    assert relerr(got, want) < 1e-14


def test_fft_real_matches_numpy_after_scramble(ct2, orc):
    """End to end in ct2's order of calls (fft_test.cpp:108-127): scramble on
    the device, then fft_real; the half-complex result equals numpy's rfft."""
    N = 1024
    x = np.random.default_rng(3).standard_normal(N)
    Xs = ct2.scramble(x, 1, N, 1)
    assert np.array_equal(Xs, orc.scramble(x, 1, N, 1))
    H = ct2.fft_real(Xs, N)
    F = np.fft.rfft(x)
    assert np.allclose(H[:N // 2 + 1], F.real, atol=1e-11)
    assert np.allclose(H[N - 1:N // 2:-1], F.imag[1:N // 2], atol=1e-11)


@pytest.mark.parametrize("N1,N2,N3", [(1, 16, 1), (3, 64, 5), (2, 1024, 4)])
def test_scramble(ct2, orc, N1, N2, N3):
    X = np.random.default_rng(N2).standard_normal(N1 * N2 * N3)
    assert np.array_equal(ct2.scramble(X, N1, N2, N3), orc.scramble(X, N1, N2, N3))


@pytest.mark.parametrize("N", [4, 16, 32])
def test_transposes(ct2, orc, N):
    X = np.random.default_rng(N).standard_normal(N ** 3)
    assert np.array_equal(ct2.transpose_yz(X, N), orc.transpose_yz(X, N))
    Y = ct2.transpose_r2c(X, N)
    assert np.array_equal(Y, orc.transpose_r2c(X, N))
    assert np.array_equal(ct2.transpose_c2r(Y, N), orc.transpose_c2r(Y, N))


@pytest.mark.parametrize("N", [8, 32, 64])
def test_fft_r2c_c2r_3d(ct2, orc, N):
    X = np.random.default_rng(N).standard_normal(N ** 3)
    Y = ct2.fft_r2c_3d(X)
    assert relerr(Y, orc.fft_r2c_3d(X.reshape(N, N, N))) < 1e-14
    R = np.fft.rfftn(X.reshape(N, N, N))  # [i][j][k<N/2+1]
    h = N * N * (N // 2 + 1)
    assert np.allclose(Y[:h], R.real.ravel(), atol=1e-10 * N ** 1.5)
    assert np.allclose(Y[h:], R.imag.ravel(), atol=1e-10 * N ** 1.5)
    Z = ct2.fft_c2r_3d(Y, N)
    assert relerr(Z, orc.fft_c2r_3d(Y, N).ravel()) < 1e-14
    assert relerr(Z / N ** 3, X) < 1e-14


def test_twiddles(ct2, orc):
    for N in (8, 1000, 4096):
        for a, b in zip(ct2.twiddles(N), orc.twiddles(N)):
            assert np.allclose(a, b, rtol=0, atol=1e-15)


def test_argument_checks(ct2):
    from ct2_amd._lib import Ct2Error
    with pytest.raises(Ct2Error):
        ct2.fft_complex(np.zeros(2 * 3000), 3000)
    with pytest.raises(Ct2Error):
        ct2.fft_real(np.zeros(8 * 8192), 8192)
    with pytest.raises(Ct2Error):
        ct2.fft_complex(np.zeros(2 * 3 * 16), 16, lanes=3)
