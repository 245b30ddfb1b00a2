"""GPU parity of the 3-D FFT against the FP64 oracle and numpy, through the C
ABI (the transform the PM path needs: fft.hpp:28-29, particle.cpp:230,261)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_fft3d_fp32_vs_oracle(ct2, orc, n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal((n,) * 3) + 1j * rng.standard_normal((n,) * 3)
    for d in (1, -1):
        got = ct2.fft3d(z.astype(np.complex64), d)
        want = orc.fft3d(z, d)
        assert rel_l2(got, want) < 1e-6, (n, d, rel_l2(got, want))


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_fft3d_fp64_vs_oracle(ct2, orc, n):
    rng = np.random.default_rng(100 + n)
    z = rng.standard_normal((n,) * 3) + 1j * rng.standard_normal((n,) * 3)
    for d in (1, -1):
        assert rel_l2(ct2.fft3d(z, d), orc.fft3d(z, d)) < 1e-14


def test_fft3d_golden(ct2):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
    for n in (8, 16):
        assert rel_l2(ct2.fft3d(g[f"c3d_in_{n}"], 1), g[f"c3d_out_{n}"]) < 1e-14
        assert rel_l2(ct2.fft3d(g[f"c3d_in_{n}"].astype(np.complex64), 1), g[f"c3d_out_{n}"]) < 1e-6


@pytest.mark.parametrize("n", [128, 256])
def test_fft3d_fp32_vs_numpy_fp64(ct2, n):
    """The GPU configuration's gate: 256^3 complex FP32, rel L2 <= 1e-5 against
    an FP64 transform, forward and round trip."""
    rng = np.random.default_rng(256)
    z = (rng.standard_normal((n,) * 3) + 1j * rng.standard_normal((n,) * 3)).astype(np.complex64)
    want = np.fft.fftn(z.astype(np.complex128))
    got = ct2.fft3d(z, 1)
    assert rel_l2(got, want) < 1e-5
    back = ct2.fft3d(got, -1) / n ** 3
    assert rel_l2(back, z.astype(np.complex128)) < 1e-5


def test_fft3d_512_round_trip_and_delta(ct2):
    """512^3 (the PM grid of the clustered and stepping configurations):
    delta -> constant, and forward/inverse round trip."""
    n = 512
    buf = ct2.DeviceBuffer(n ** 3 * 8)
    z = np.zeros((n,) * 3, np.complex64)
    z[3, 5, 7] = 1.0
    buf.upload(z)
    ct2.fft3d_device(buf, n, 1)
    out = buf.download(np.empty_like(z))
    k = np.arange(n)
    ph = np.exp(-2j * np.pi * 7 * k / n)
    assert np.allclose(out[0, 0, :], ph, atol=1e-5)
    assert np.allclose(np.abs(out[::37, ::41, ::43]), 1.0, atol=1e-5)
    rng = np.random.default_rng(512)
    z = (rng.standard_normal((n,) * 3, dtype=np.float32) + 1j * rng.standard_normal((n,) * 3, dtype=np.float32)).astype(np.complex64)
    buf.upload(z)
    ct2.fft3d_device(buf, n, 1)
    ct2.fft3d_device(buf, n, -1)
    back = buf.download(np.empty_like(z)) / np.float32(n ** 3)
    assert rel_l2(back, z) < 1e-5
    buf.close()


def test_fft3d_rejects_bad_sizes(ct2):
    from ct2_amd import Ct2Error
    buf = ct2.DeviceBuffer(1024)
    for n, d in ((12, 1), (4, 1), (2048, 1), (64, 0)):
        with pytest.raises(Ct2Error):
            ct2.fft3d_device(buf, n, d)
    buf.close()


def test_cfg2_fft_of_the_cic_density(ct2):
    """The GPU configuration's FFT gate on the field it names: the CIC density
    of N = 128^3 lattice + Gaussian(0.2 cell) particles on the 256^3 grid (the
    library's own deposit), as complex FP32, transformed forward and back on
    the device; rel L2 <= 1e-5 against numpy's FP64 transform of the same
    values, for the density and for its contrast (the mean's k = 0 term
    removed, so that the check is not carried by that one mode)."""
    side, ng = 128, 256
    x, y, z = ct2.ic_lattice(side, 0.2, 1)
    ct2.particle_allocate(x.size)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(float(np.float32(1.0 / (50 * side))), float(np.float32(1.0 / x.size)), 0)
    ct2.pm_assignment(1)
    ct2.pm_init(ng, 1.25)
    ct2.pm_forces()
    W = ct2.pm_density_download()
    rho = (W.astype(np.float64) * 2.0 ** -32).astype(np.float32)  # particles per cell
    assert rho.sum(dtype=np.float64) == pytest.approx(x.size, rel=1e-6)
    for field in (rho, rho / np.float32(rho.mean()) - np.float32(1.0)):
        zc = field.astype(np.complex64)
        want = np.fft.fftn(zc.astype(np.complex128))
        got = ct2.fft3d(zc, 1)
        assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
        back = ct2.fft3d(got, -1) / ng ** 3
        assert rel_l2(back, zc.astype(np.complex128)) < 1e-5
