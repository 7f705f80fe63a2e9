"""Post-processing of the accuracy studies (CPU only)."""
import math

import numpy as np

from lettuce_amd.studies import _strouhal


def test_strouhal_of_a_noisy_sine():
    """St = f D / U: a lift signal of known period, sampled every 50 steps,
    with noise and a faster acoustic component, gives St to 0.5 %"""
    every, diameter, u = 50, 64, 0.05
    for st in (0.192, 0.2113, 0.165):
        period_steps = diameter / (u * st)
        t = np.arange(800) * every
        rng = np.random.default_rng(1)
        cl = (0.12 * np.sin(2 * np.pi * t / period_steps + 0.3) + 0.03
              + 0.02 * rng.standard_normal(t.size) + 0.01 * np.sin(2 * np.pi * t / 370.0))
        assert abs(_strouhal(cl, every, diameter, u) / st - 1) < 5e-3


def test_strouhal_degenerate_signals():
    assert math.isnan(_strouhal([0.0] * 100, 50, 64, 0.05))
    assert math.isnan(_strouhal([1.0, 2.0, 3.0], 50, 64, 0.05))


def test_spectral_velocity_is_solenoidal_with_the_requested_spectrum():
    import lettuce_amd as lt
    n = 256
    u = lt.spectral_velocity_2d(n, seed=3, k0=4.0, u_rms=0.04)
    assert u.shape == (2, n, n)
    assert abs(np.sqrt((u * u).sum(0).mean()) - 0.04) < 1e-15
    U = np.fft.fft2(u)
    k = np.fft.fftfreq(n, 1.0 / n)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    assert np.abs(kx * U[0] + ky * U[1]).max() < 1e-10 * np.abs(U).max()
    kk = np.rint(np.hypot(kx, ky)).astype(int)
    shell = np.bincount(kk.ravel(), (np.abs(U) ** 2).sum(0).ravel())
    assert 3 <= int(np.argmax(shell)) <= 5
    # same seed, same field; another seed, another field
    assert np.array_equal(u, lt.spectral_velocity_2d(n, seed=3))
    assert not np.allclose(u, lt.spectral_velocity_2d(n, seed=4))


def test_spectral_downsample_samples_band_limited_fields_and_cuts_high_modes():
    import torch
    import lettuce_amd as lt
    u = lt.spectral_velocity_2d(128, seed=0)
    c = lt.spectral_downsample(torch.from_numpy(u))
    assert c.shape == (2, 64, 64)
    assert np.abs(c.numpy() - u[:, ::2, ::2]).max() < 1e-15
    x = np.arange(128)
    high = np.cos(2 * np.pi * 40 * x / 128)[:, None] * np.ones((1, 128))
    low = np.cos(2 * np.pi * 5 * x / 128)[:, None] * np.ones((1, 128))
    c = lt.spectral_downsample(torch.from_numpy(high + low)).numpy()
    assert np.abs(c - low[::2, ::2]).max() < 1e-13
