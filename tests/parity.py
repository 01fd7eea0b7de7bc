"""Parity metrics shared by the tests (per-particle force error with a floor,
global relative L2, energy relative error with an absolute floor)."""
import numpy as np

FORCE_RTOL = 1e-5     # per-particle ||dF|| / max(||F||, 1e-6 max||F||)
ENERGY_RTOL = 1e-5


def force_errors(F, Fref, mask):
    F = np.asarray(F, np.float64)[mask]
    R = np.asarray(Fref, np.float64)[mask]
    d = np.linalg.norm(F - R, axis=1)
    nr = np.linalg.norm(R, axis=1)
    if nr.size == 0 or nr.max() == 0.0:
        return float(d.max(initial=0.0)), 0.0
    floor = 1e-6 * nr.max()
    per = d / np.maximum(nr, floor)
    glob = np.linalg.norm(F - R) / np.linalg.norm(R)
    return float(per.max()), float(glob)


def energy_error(E, Eref, floor):
    """|E - Eref| / max(|Eref|, floor)."""
    return abs(float(E) - float(Eref)) / max(abs(float(Eref)), floor)
