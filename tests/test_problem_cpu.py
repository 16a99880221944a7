"""CPU checks of the closed-form fields used as solve references."""
import numpy as np

from fmm3dbie_amd import problem


def test_mie_total_field_vanishes_on_sphere():
    d = problem.plane_wave_dirs(4, 2)
    x = problem.plane_wave_dirs(30, 9)
    for k in (1.1, 5.0):
        assert np.abs(problem.mie_sound_soft(x, d, k) + problem.plane_waves(x, d, k)).max() < 1e-13


def test_mie_radiates():
    """Outgoing: r e^{-ikr} u_s tends to a far-field pattern (Sommerfeld condition)."""
    d = problem.plane_wave_dirs(1, 2)
    xh = problem.plane_wave_dirs(5, 3)
    k = 2.0
    f = [r * np.exp(-1j * k * r) * problem.mie_sound_soft(r * xh, d, k)[0] for r in (400.0, 800.0)]
    assert np.abs(f[0] - f[1]).max() < 5e-3 * np.abs(f[1]).max()
