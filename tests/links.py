"""Shared link configurations and seeded waveforms for the tests (inputs
built with the oracle's generators)."""
import numpy as np

import oracle as O

C_LIGHT = 299792458.0
WAVELENGTH = 1.55e-6
F_C = C_LIGHT / WAVELENGTH


def ssmf(length=80e3, alpha_db_km=0.2, beta2_ps2_km=-21.7, gamma=1.3e-3, kappa=1.0):
    """SSMF in SI units: alpha = db_to_natural(dB/m) (functions.cpp:90-92)."""
    return dict(alpha=O.db_to_natural(alpha_db_km * 1e-3), beta2=beta2_ps2_km * 1e-27,
                gamma=gamma, length=length, wavelength=WAVELENGTH, kappa=kappa)


def qam_waveform(seed, n_samples, sps=2, power_dbm=2.0, roll_off=0.1, modes=1, baud=32e9):
    """Seeded 16QAM, RRC-shaped at `sps` samples/symbol, scaled to the launch
    power per mode.  Returns (field, symbols, dt)."""
    n_sym = n_samples // sps
    syms, _ = O.qam16_symbols(seed, n_sym * modes)
    syms = syms * np.sqrt(O.dbm_to_watts(power_dbm))
    x = O.rrc_shaping(syms, sps, roll_off, modes=modes)
    return x, syms, 1.0 / (baud * sps)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
