"""Parity tolerances, in one place.

CAPE/CIN: FP32 results within 0.5 J/kg or 0.1 % of the reference, whichever
is larger.  Bunkers motion within 0.05 m/s, SRH within 1 m^2/s^2.  Level
indices exact apart from ties (theta-e equal to a few ulp).
"""
import numpy as np

CAPE_ABS = 0.5
CAPE_REL = 1e-3
BUNKERS_ABS = 0.05
SRH_ABS = 1.0


def cape_tol(ref):
    return np.maximum(CAPE_ABS, CAPE_REL * np.abs(np.asarray(ref, np.float64)))


def cape_ok(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(got - ref) <= cape_tol(ref)
