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


PERTURB_SEEDS = range(1, 17)
# how far a converged sub-step's accepted iterate may move in the oracle's
# iterate-perturbed runs [K]: a saturated sub-step accepts iterates within
# about 2e-4 / |1 - f'| <= 7e-5 K of the fixed point (f' <= -2), and the
# test filters the ones outside the band
ITERATE_AMP = 1e-4


def ill_conditioned_and_matched(run, cols, gots):
    """For columns `cols` where two device results disagree beyond the
    tolerance: the oracle's FP32 run, its FP64 run, 16 runs with every exp,
    log and pow moved by up to an ulp and 16 that also accept another point
    of each sub-step's convergence band (run(cols, fp64=..., perturb=...,
    perturb_iterate=...) -> oracle outputs).  Returns (ill, matched): ill[j]
    when those runs themselves disagree beyond the tolerance on CAPE or CIN
    (the reference's result there turns on the last bit of its C library or
    on which iterate of its tolerance band it accepts), matched[j] when every
    device result in `gots` ((cape, cin) pairs over cols) is within the
    tolerance of at least one of the runs."""
    runs = [run(cols), run(cols, fp64=True)]
    runs += [run(cols, perturb=s) for s in PERTURB_SEEDS]
    runs += [run(cols, perturb=s, perturb_iterate=ITERATE_AMP) for s in PERTURB_SEEDS]
    base = runs[0]
    ill = np.zeros(len(cols), bool)
    for o in runs[1:]:
        ill |= ~(cape_ok(o[0], base[0]) & cape_ok(o[1], base[1]))
    matched = np.ones(len(cols), bool)
    for cape, cin in gots:
        m = np.zeros(len(cols), bool)
        for o in runs:
            m |= cape_ok(cape, o[0]) & cape_ok(cin, o[1])
        matched &= m
    return ill, matched
