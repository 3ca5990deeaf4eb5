"""Sampled spectra against the reference's own golden vector
(test/test_spectrum.cpp:256-278): Sampled([400, 500, 600] nm, [0.1, 0.2, 0])
resampled on 350..650 nm in 25 nm steps gives
[0, 0, 0.1, 0.125, 0.15, 0.175, 0.2, 0.15, 0.1, 0.05, 0, 0, 0] to 1e-6.
The same values come out of the oracle's spectrum evaluation (the
renderer's per-wavelength lookup, Spectral::fromSampled in Single mode,
src/liar/kernel/spectral.cpp:185-210) and of Sampled.__call__
(sampled.cpp:102-123); the GPU's lookup is checked against them in
tests/test_shading_gpu.py."""
import numpy as np
import pytest

import oracle
from liar_amd import spectra

WS_M = [350e-9, 375e-9, 400e-9, 425e-9, 450e-9, 475e-9, 500e-9, 525e-9, 550e-9, 575e-9, 600e-9, 625e-9, 650e-9]
VS = [0.0, 0.0, 0.1, 0.125, 0.15, 0.175, 0.2, 0.15, 0.1, 0.05, 0.0, 0.0, 0.0]


def golden():
    return spectra.Sampled([400e-9, 500e-9, 600e-9], [0.1, 0.2, 0.0])


def test_resample_matches_reference_vector():
    r = golden().resample(WS_M)
    np.testing.assert_allclose(r.wavelengths * 1e-9, WS_M, atol=1e-15)
    np.testing.assert_allclose(r.values, VS, atol=1e-6)


def test_call_matches_reference_vector():
    s = golden()
    np.testing.assert_allclose([s(w) for w in WS_M], VS, atol=1e-6)
    # doCall returns the last value at the last wavelength
    assert spectra.Sampled([400, 500], [0.3, 0.7])(500) == pytest.approx(0.7)


def test_oracle_lookup_matches_reference_vector():
    got = oracle.spectrum_eval(golden(), np.array(WS_M) * 1e9)
    np.testing.assert_allclose(got, VS, atol=1e-6)
    # fromSampled is 0 at and beyond the last wavelength, linear inside
    s = spectra.Sampled([400, 500], [0.3, 0.7])
    np.testing.assert_allclose(oracle.spectrum_eval(s, [399.9, 400, 450, 499.9, 500, 510]),
                               [0, 0.3, 0.5, 0.7 - 0.4e-3, 0, 0], atol=1e-6)


def test_reflectant_normalisation():
    """Spectral::fromSampled divides a Reflectant by the table maximum when
    it exceeds 1 (spectral.cpp:200-207)."""
    s = spectra.Sampled([400, 500, 600], [1.0, 4.0, 2.0])
    np.testing.assert_allclose(oracle.spectrum_eval(s, [450, 550], reflectant=True), [2.5 / 4, 3.0 / 4], rtol=1e-7)
    np.testing.assert_allclose(oracle.spectrum_eval(s, [450, 550], reflectant=False), [2.5, 3.0], rtol=1e-7)
