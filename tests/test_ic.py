"""CPU checks of the seeded initial states (host code of libmhd2d.so)."""
import os

import numpy as np
import pytest

import mhd2dcfdfftw_open_amd as m
import oracle_lib as orc

pytestmark = pytest.mark.skipif(not os.path.exists(m.LIB_PATH), reason="libmhd2d.so not built")


def test_orszag_tang_values():
    nx, ny = 16, 8
    u = m.orszag_tang(nx, ny)
    x = np.arange(nx)[:, None] * 2 * np.pi / nx
    y = np.arange(ny)[None, :] * 2 * np.pi / ny
    s = 1 / np.sqrt(4 * np.pi)
    np.testing.assert_allclose(u[0], 25 / (36 * np.pi))
    np.testing.assert_allclose(u[7], 5 / (12 * np.pi))
    np.testing.assert_allclose(u[1], -np.sin(y) + 0 * x, atol=1e-15)
    np.testing.assert_allclose(u[2], np.sin(x) + 0 * y, atol=1e-15)
    np.testing.assert_allclose(u[4], -np.sin(y) * s + 0 * x, atol=1e-15)
    np.testing.assert_allclose(u[5], np.sin(2 * x) * s + 0 * y, atol=1e-15)


def test_turbulence_rms_spectrum_and_divb():
    nx, ny = 128, 64
    u = m.turbulence(nx, ny, seed=42, rms=0.1, kmin=1, kmax=16)
    np.testing.assert_allclose(np.sqrt(np.mean(np.sum(u[1:4] ** 2, axis=0))), 0.1, rtol=1e-12)
    dB = u[4:7].copy()
    dB[0] -= 1.0
    np.testing.assert_allclose(np.sqrt(np.mean(np.sum(dB ** 2, axis=0))), 0.1, rtol=1e-12)
    assert np.all(u[0] == 1.0) and np.all(u[7] == 1.0)
    # spectral div B of the in-plane field vanishes (dB from A_z)
    kx = np.fft.fftfreq(nx, 1 / nx)[:, None]
    ky = np.fft.fftfreq(ny, 1 / ny)[None, :]
    div = kx * np.fft.fft2(u[4]) + ky * np.fft.fft2(u[5])
    assert np.abs(div).max() < 1e-10 * np.abs(np.fft.fft2(u[5])).max()
    # only 1 <= |k| <= 16 is populated
    F = np.abs(np.fft.fft2(u[3]))
    k = np.sqrt(kx ** 2 + ky ** 2)
    assert F[(k < 1) | (k > 16)].max() < 1e-10 * F.max()


def test_turbulence_seeded_and_deterministic():
    a = m.turbulence(32, 32, seed=1)
    b = m.turbulence(32, 32, seed=1)
    c = m.turbulence(32, 32, seed=2)
    np.testing.assert_array_equal(a, b)
    assert np.abs(a - c).max() > 1e-3


def test_linear_wave_is_alfven_eigenmode():
    u = m.linear_wave(32, 32, seed=3, amp=1e-4)
    np.testing.assert_array_equal(u[3], -u[6])
    assert np.abs(u[6]).max() == pytest.approx(1e-4, rel=1e-2)


def test_oracle_energy_of_turbulence_ic():
    u = m.turbulence(64, 64, seed=5, rms=0.1)
    d = 2 * np.pi / 64
    e = orc.energy(u, d, d)
    # 1/2 <rho u^2> + p/(g-1) + 1/2 <B^2> with <B^2> = 1 + rms^2
    want = 4 * np.pi ** 2 * (0.5 * 0.01 + 1.5 + 0.5 * (1 + 0.01))
    assert abs(e - want) < 1e-12 * want
