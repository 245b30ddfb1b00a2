"""Writes fft_golden.npz: seeded FP64 inputs and numpy.fft (pocketfft) outputs.

The reference pins its FFTs only against FFTW at run time (fft_test.cpp:14-79,
N = 8..4096, forward, unnormalised, FFTW_FORWARD sign); FFTW is absent here,
so numpy's FP64 transforms with the same sign and normalisation stand in.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20230803)
    out = {}
    for N in (8, 16, 32, 64, 128, 256, 1024, 4096):
        x = rng.uniform(size=N) + 1j * rng.uniform(size=N)  # rand1()-like inputs, fft_test.cpp:84-87
        out[f"c1d_in_{N}"] = x
        out[f"c1d_out_{N}"] = np.fft.fft(x)
    for N in (8, 32, 128, 512, 4096):
        x = rng.uniform(size=N)
        out[f"r1d_in_{N}"] = x
        out[f"r1d_out_{N}"] = np.fft.rfft(x)
    for n in (8, 16):
        z = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
        out[f"c3d_in_{n}"] = z
        out[f"c3d_out_{n}"] = np.fft.fftn(z)
    r = rng.standard_normal((8, 8, 8))
    out["r3d_in_8"] = r
    out["r3d_out_8"] = np.fft.rfftn(r)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
