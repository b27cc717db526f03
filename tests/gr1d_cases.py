"""A synthetic field-line profile for the 1D GR gap (the datasets of the
reference's coef.h5: r, Bu1, Bu3, rho, ju1, th, dpsidth): smooth
functions over r in [1.15, 12], so the tests need no reference file.  The
grid xi in [-1.2, -0.2] maps to r in about [1.8, 6] for a = 0.99."""
import numpy as np

N, GUARD, LOWER, SIZE = 256, 3, -1.2, 1.0


def profile(n=400):
    r = np.linspace(1.15, 12.0, n)
    return {"r": r, "Bu1": 1.0 / r ** 3 + 0.01, "Bu3": 0.3 / r ** 2,
            "rho": 0.2 * np.sin(r), "ju1": 0.05 * np.cos(r), "th": 0.3 + 0.02 * r,
            "dpsidth": 1.0 + 0.1 * r}
