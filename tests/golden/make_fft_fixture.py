"""Writes fft_8x16x32.npz: a seeded complex (8,16,32) array with its forward
(exp(-2 pi i ...)) and unnormalised inverse 3-D DFT as computed by numpy's
pocketfft.  The fixture pins the oracle's FFT, which stands in for the
reference's absent Sandia fft_3d / FFTW 2.1.5 (poisson.cpp:87,431,453)."""
import os

import numpy as np

rng = np.random.default_rng(20240)
a = rng.standard_normal((8, 16, 32)) + 1j * rng.standard_normal((8, 16, 32))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_8x16x32.npz")
np.savez_compressed(out, input=a, forward=np.fft.fftn(a), inverse_unnormalised=np.fft.ifftn(a) * a.size)
print("wrote", out)
