"""Pins the FP64 FFT oracle: golden vectors (numpy FP64 in place of the
reference's FFTW, fft_test.cpp:14-79), a brute-force DFT, and the layout,
ordering and inverse conventions of src/fft/*."""
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def bitrev_perm(N):
    b = N.bit_length() - 1
    return np.array([int(format(i, f"0{b}b")[::-1], 2) if b else 0 for i in range(N)])


def to_half_complex(F, N):
    """fft_test.cpp:143-155: X[k] = Re F_k, X[N-k] = Im F_k, 0<k<N/2."""
    hc = np.empty(N)
    hc[0] = F[0].real
    hc[N // 2] = F[N // 2].real
    for k in range(1, N // 2):
        hc[k] = F[k].real
        hc[N - k] = F[k].imag
    return hc


def brute_dft(x, sign=-1):
    N = x.size
    n = np.arange(N)
    return np.exp(sign * 2j * np.pi * np.outer(n, n) / N) @ x


@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 256, 1024, 4096])
def test_complex_forward_golden(orc, N):
    x, ref = GOLD[f"c1d_in_{N}"], GOLD[f"c1d_out_{N}"]
    p = bitrev_perm(N)
    X = np.empty(2 * N)
    X[0::2], X[1::2] = x[p].real, x[p].imag  # scramble(X, 1, N, 2) as fft_test.cpp:88
    Y = orc.fft_complex(X, N, 1, 1)
    y = Y[0::2] + 1j * Y[1::2]
    assert np.max(np.abs(y - ref)) <= 1e-13 * N * np.max(np.abs(ref))


@pytest.mark.parametrize("N", [8, 32, 128, 512, 4096])
def test_real_forward_golden(orc, N):
    x, ref = GOLD[f"r1d_in_{N}"], GOLD[f"r1d_out_{N}"]
    X = orc.scramble(x, 1, N, 1)  # fft_test.cpp:139
    Y = orc.fft_real(X, N, 1, 1)
    hc = to_half_complex(ref, N)
    assert np.max(np.abs(Y - hc)) <= 1e-13 * N * np.max(np.abs(hc))


def test_four_lane_layouts(orc):
    """fft_test.cpp:82-97 / 104-110: complex<simd_f64> is {re[4], im[4]},
    simd_f64 is [n][4]; every lane is an independent transform."""
    rng = np.random.default_rng(3)
    N = 64
    z = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    X = np.empty((N, 8))
    X[:, :4], X[:, 4:] = z.real, z.imag
    Xs = orc.scramble(X.ravel(), 1, N, 8).reshape(N, 8)
    Y = orc.fft_complex(Xs.ravel(), N, 1, 4).reshape(N, 8)
    assert np.allclose(Y[:, :4] + 1j * Y[:, 4:], np.fft.fft(z, axis=0), rtol=0, atol=1e-12)
    r = rng.standard_normal((N, 4))
    R = orc.fft_real(orc.scramble(r.ravel(), 1, N, 4), N, 1, 4).reshape(N, 4)
    for lane in range(4):
        assert np.allclose(R[:, lane], to_half_complex(np.fft.rfft(r[:, lane]), N), atol=1e-12)


@pytest.mark.parametrize("N", [2, 4, 8, 16, 32, 64, 128])
def test_complex_matches_brute_force_both_signs(orc, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    p = bitrev_perm(N)
    X = np.empty(2 * N)
    X[0::2], X[1::2] = x[p].real, x[p].imag
    for d, sign in ((1, -1), (-1, 1)):
        Y = orc.fft_complex(X, N, d, 1)
        assert np.allclose(Y[0::2] + 1j * Y[1::2], brute_dft(x, sign), atol=1e-12 * N)


def test_inverse_round_trip_and_conjugation(orc):
    """dir=-1 is the true inverse (intent; fft_complex_vector.cpp:57-58 is
    wrong): fwd then inv gives N x, and inv(x) = conj(fwd(conj x))."""
    N = 256
    rng = np.random.default_rng(7)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    p = bitrev_perm(N)

    def run(v, d):
        X = np.empty(2 * N)
        X[0::2], X[1::2] = v[p].real, v[p].imag
        Y = orc.fft_complex(X, N, d, 1)
        return Y[0::2] + 1j * Y[1::2]

    f = run(x, 1)
    assert np.allclose(run(f, -1), N * x, atol=1e-10)
    assert np.allclose(run(x, -1), np.conj(run(np.conj(x), 1)), atol=1e-12)


def test_invariants_delta_parseval_shift(orc):
    N = 128
    p = bitrev_perm(N)

    def fwd(v):
        X = np.empty(2 * N)
        X[0::2], X[1::2] = v[p].real, v[p].imag
        Y = orc.fft_complex(X, N, 1, 1)
        return Y[0::2] + 1j * Y[1::2]

    d = np.zeros(N, complex)
    d[0] = 1
    assert np.allclose(fwd(d), 1.0, atol=1e-15)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    F = fwd(x)
    assert np.isclose(np.sum(np.abs(F) ** 2), N * np.sum(np.abs(x) ** 2), rtol=1e-13)
    k = np.arange(N)
    assert np.allclose(fwd(np.roll(x, 5)), F * np.exp(-2j * np.pi * 5 * k / N), atol=1e-11)


def test_twiddles_closed_form(orc):
    N = 64
    c, s, i = orc.twiddles(N)
    n = np.arange(N)
    assert np.allclose(c, np.cos(2 * np.pi * n / N), atol=1e-15)
    assert np.allclose(s, -np.sin(2 * np.pi * n / N), atol=1e-15)
    assert np.allclose(i, np.sin(2 * np.pi * n / N), atol=1e-15)


def test_scramble_is_block_bit_reversal(orc):
    """scramble(X, N1, N2, N3): each of N1 groups permutes N2 blocks of N3."""
    N1, N2, N3 = 3, 16, 5
    X = np.arange(N1 * N2 * N3, dtype=float)
    Y = orc.scramble(X, N1, N2, N3).reshape(N1, N2, N3)
    p = bitrev_perm(N2)
    assert np.array_equal(Y, X.reshape(N1, N2, N3)[:, p, :])
    assert np.array_equal(orc.scramble(Y.ravel(), N1, N2, N3), X)  # an involution


@pytest.mark.parametrize("n", [8, 16])
def test_fft3d_golden(orc, n):
    z, ref = GOLD[f"c3d_in_{n}"], GOLD[f"c3d_out_{n}"]
    assert np.max(np.abs(orc.fft3d(z, 1) - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.allclose(orc.fft3d(ref, -1), n ** 3 * z, atol=1e-9)


def test_fft3d_32_cubed_against_numpy(orc):
    """The 32^3 complex FP64 transform of the CPU configuration."""
    rng = np.random.default_rng(32)
    z = rng.standard_normal((32,) * 3) + 1j * rng.standard_normal((32,) * 3)
    assert np.allclose(orc.fft3d(z, 1), np.fft.fftn(z), atol=1e-10)


def test_r2c_c2r_3d_layout(orc):
    """fft_r2c_3d intent: planar FFTW half spectrum [N][N][N/2+1] with the
    imaginary plane at N^2(N/2+1) (transpose.cpp:36, fft_test.cpp:176-185)."""
    r, ref = GOLD["r3d_in_8"], GOLD["r3d_out_8"]
    N, h = 8, 5
    out = orc.fft_r2c_3d(r)
    assert np.allclose(out[: N * N * h], ref.real.ravel(), atol=1e-12)
    assert np.allclose(out[N * N * h:], ref.imag.ravel(), atol=1e-12)
    assert np.allclose(orc.fft_c2r_3d(out, N), N ** 3 * r, atol=1e-10)


def test_transposes_intended_maps(orc):
    N = 8
    rng = np.random.default_rng(5)
    X = rng.standard_normal(N ** 3)
    T = orc.transpose_yz(X, N).reshape(N, N, N)
    assert np.array_equal(T, X.reshape(N, N, N).transpose(0, 2, 1))
    Y = orc.transpose_r2c(X, N)
    h, dimag = N // 2 + 1, N * N * (N // 2 + 1)
    Xc = X.reshape(N, N, N)
    # out[ix][iz][iy] = in[ix][iy][iz] (real), in[ix][N-iy][iz] (imag, 0<iy<N/2)
    re = Y[:dimag].reshape(N, N, h)
    im = Y[dimag:].reshape(N, N, h)
    for iy in range(h):
        assert np.array_equal(re[:, :, iy], Xc[:, iy, :])
        if 0 < iy < N // 2:
            assert np.array_equal(im[:, :, iy], Xc[:, N - iy, :])
        else:
            assert np.all(im[:, :, iy] == 0)
    assert np.array_equal(orc.transpose_c2r(Y, N), X)
