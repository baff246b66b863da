"""Soundings that drive getcape into the branches no reference fixture reaches.

One 128-column, 30-level sigma grid (the synthetic generator, seed 5) with:

* every 7th column below freezing at the surface (the ``ts > 0`` gate,
  CAPE_CODE_model_lev.f90:77, zeros);
* every 11th column from column 3 with its surface at 450 hPa and every
  level above 440 hPa at most: most-unstable then keeps the surface parcel
  and MUlev -999999 (:264-267), and with surface temperatures of 19-31 degC
  at 450 hPa the moist-adiabat solve of columns 47, 102 (and 124 for the
  reversible adiabats) exceeds the 100-iteration limit, which zeroes CAPE and
  CIN (:464-474);
* a NaN temperature at level 10 of column 5 and a NaN dewpoint at level 20
  of column 6, which end those columns' ascents through the NaN comparisons
  of the solve (:475) and the trapezoids.

With ``ml_depth`` 1 m the second level lies above the mixed layer (:287-293);
with 1e6 m the whole column lies inside it (:295-301).  The oracle matches the
reference Fortran bit for bit on all of these (tests/test_oracle.py).
"""
import numpy as np

from xcape_amd import synthetic

# (source, ml_depth) pairs that select the branches
CASES = ((3, 1.0), (3, 1.0e6), (3, 500.), (2, 500.), (1, 500.))


def grid():
    c = synthetic.sigma_columns(128, 30, 5)
    p, t, td, ps, ts, tds = (c[k].copy() for k in ('p', 't', 'td', 'ps', 'ts', 'tds'))
    ts[::7] = -2.0
    ps[3::11] = 450.0
    p[3::11] = np.minimum(p[3::11], 440.0)
    t[5, 10] = np.nan
    td[6, 20] = np.nan
    return p, t, td, ps, ts, tds
