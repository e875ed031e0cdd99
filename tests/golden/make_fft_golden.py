"""Writes fft_golden.npz: seeded real fields and their unnormalised 3-D r2c
spectra computed with numpy.fft.rfftn (the library routine the FFTW r2c of
include/solver.h:221 must agree with).  Run from the repo root:
    python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(20241202)
out = {}
for shape in [(8, 6, 4, 1), (6, 10, 12, 3), (5, 9, 10, 3), (16, 8, 6, 1)]:
    tag = "x".join(map(str, shape))
    x = rng.standard_normal(shape)
    out["x_" + tag] = x
    out["X_" + tag] = np.fft.rfftn(x, axes=(0, 1, 2))
np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
