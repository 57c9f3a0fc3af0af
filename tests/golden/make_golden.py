"""Writes the golden fixtures of tests/golden from the reference's own
sources, run in the container where the reference tree is mounted:

  fft_roundtrip_2x2x2.json  the data of the reference's FFT test
                            (test/fftAdaptersTest.cpp:6-22: input 1..8,
                            forward + inverse must return the input within
                            1e-6) plus the forward spectrum kissfft gives
  kissfft_8x16x32.npz       a seeded complex64 array, its forward transform
                            and inverse round trip through the reference's
                            vendored kissfft (libs/kissfft via
                            include/kissFFTAdapter.h), built from the
                            reference tree into oracle/_ref by oracle/Makefile

    make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    if not O.kissfft_available():
        raise SystemExit("oracle/_ref/libkissfft.so missing: run `make -C oracle` where the reference is mounted")
    a = np.arange(1, 9, dtype=np.complex64).reshape(2, 2, 2)
    f = O.kiss_fftnd(a)
    b = O.kiss_fftnd(f, inverse=True)
    assert np.max(np.abs(b - a)) <= 1e-6
    with open(os.path.join(HERE, "fft_roundtrip_2x2x2.json"), "w") as fh:
        json.dump({"source": "test/fftAdaptersTest.cpp:6-22 (input, tolerance); forward by kissfft",
                   "dims": [2, 2, 2], "input": [float(v.real) for v in a.ravel()],
                   "forward_re": [float(v.real) for v in f.ravel()],
                   "forward_im": [float(v.imag) for v in f.ravel()],
                   "tolerance": 1e-6}, fh, indent=1)
    rng = np.random.default_rng(2024)
    x = (rng.standard_normal((8, 16, 32)) + 1j * rng.standard_normal((8, 16, 32))).astype(np.complex64)
    fx = O.kiss_fftnd(x)
    rx = O.kiss_fftnd(fx, inverse=True)
    np.savez_compressed(os.path.join(HERE, "kissfft_8x16x32.npz"), input=x, forward=fx, inverse_roundtrip=rx)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
