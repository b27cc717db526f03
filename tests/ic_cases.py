"""Soft-photon spectra and energy ranges shared by the inverse-Compton
oracle and GPU tests: (kind, parameters, emin, emax, n0).  Kinds follow
src/radiation/spectra.h: 0 hard power law (delta, emin, emax), 1 soft power
law (alpha, emin, emax), 2 broken power law (alpha, delta, epeak, emin,
emax), 3 black body (kT), 4 mono-energetic (e0, de)."""
import numpy as np

F32_1EM5 = float(np.float32(1.0e-5))

SPECTRA = {
    "soft_pl": (1, [1.0, 1e-6, 1e-4], 1e-6, 1e-4, 1.0),
    "hard_pl": (0, [0.5, 1e-7, 1e-3], 1e-7, 1e-3, 1.0),
    # tests/test_ic.cu:31-34: broken_power_law(1.25, 1.1, e_min = 1e-5f,
    # 1e-10, 0.1), n0 = 1.5e24 / ic_path with ic_path = 0.01
    "test_ic_cu": (2, [1.25, 1.1, F32_1EM5, 1e-10, 0.1], 1e-10, 0.1, 1.50e24 / 0.01),
    "black_body": (3, [1e-6], 1e-16, 1e-3, 1.0),
    "mono": (4, [1e-5, 1e-6], 1e-7, 1e-3, 1.0),
}

THOMSON = 8.0 * np.pi / 3.0 * 7.91402e-26  # sigma_T = (8 pi / 3) r_e^2
