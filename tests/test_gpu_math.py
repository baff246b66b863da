"""The accurate mode's device arithmetic, checked on the device.

The kernels divide with the raw IEEE division sequence (the compiler's own
steps without its operand scaling), divide by per-call constants with a
correctly rounded reciprocal, and take exp and pow from short kernels
(phys.h).  Their claims, checked here elementwise through xcape_math_check:

* the raw division equals the compiler's IEEE division bit for bit wherever
  its operands and quotient are normal numbers (checked from 1e-30 to 1e30,
  both signs, far beyond the physical range), and gives NaN for a zero,
  infinite or NaN operand (which the fixed-point evaluation detects and
  re-evaluates with IEEE division);
* division by a constant equals IEEE division for the constants the kernels
  use (the ice ramp's 40 K, pressure increments) -- the CPU check of the
  same sequence over every float is in the notes of phys.h;
* the exp kernel equals the device library's expf bit for bit on (-80, 80);
* pow stays within the device library's powf accuracy: its error against the
  exactly rounded x^y (from float64 on the host) is at most 1 ulp over the
  Exner function's and theta-e's ranges, and no worse than powf's;
* the helicity chain's FP64 log and exp (xcape_math_check_f64) are within an
  ulp of ln x and e^x (long double on the host) over the doubles' whole
  range, the heights' arguments (level pressure ratios, dewpoints in K, the
  saturation vapour pressure's exponent) included, and give the C library's
  special values; its raw FP64 division equals IEEE division bit for bit
  for operands from 1e-150 to 1e150, and gives NaN for a zero, infinite or
  NaN operand.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def dev():
    import torch

    from xcape_amd import _hip
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return torch, _hip


def _bits(x):
    return np.asarray(x, np.float32).view(np.int32)


def _ulp_err(got, exact):
    """|got - exact| in ulps of the float32 nearest to the exact value."""
    got = np.asarray(got, np.float64)
    ref32 = exact.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got - exact) / ulp


def test_raw_division_is_ieee(dev):
    torch, H = dev
    rng = np.random.default_rng(11)
    n = 1 << 22
    mag = lambda: 10.0 ** rng.uniform(-30, 30, n)  # noqa: E731
    sgn = lambda: rng.choice([-1.0, 1.0], n)  # noqa: E731
    a = (sgn() * mag()).astype(np.float32)
    b = (sgn() * mag()).astype(np.float32)
    # plus the arithmetic's own scales: quotients near 1, physical magnitudes
    a[: n // 4] = rng.uniform(-400, 400, n // 4).astype(np.float32)
    b[: n // 4] = rng.uniform(1, 1.2e5, n // 4).astype(np.float32)
    q = np.abs(a.astype(np.float64) / b.astype(np.float64))
    normal = (q > 1e-37) & (q < 1e37)
    a, b = a[normal], b[normal]
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    ieee = H.math_check(1, ta, tb).cpu().numpy()
    raw = H.math_check(0, ta, tb).cpu().numpy()
    assert np.array_equal(_bits(ieee), _bits(a / b))
    mism = np.count_nonzero(_bits(raw) != _bits(ieee))
    print(f'raw division: {mism} of {a.size} differ from IEEE', flush=True)
    assert mism == 0


def test_raw_division_special_operands_give_nan(dev):
    torch, H = dev
    a = np.array([1.0, 0.0, np.inf, 2.0, np.nan, 3.0, -np.inf], np.float32)
    b = np.array([0.0, 0.0, 2.0, np.inf, 1.0, np.nan, -np.inf], np.float32)
    raw = H.math_check(0, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()).cpu().numpy()
    assert np.isnan(raw).all(), raw


@pytest.mark.parametrize('c', [40.0, 500.0, 100.0, 1000.0, 2500.0, 1.0, 37.5])
def test_constant_division_is_ieee(dev, c):
    torch, H = dev
    rng = np.random.default_rng(int(c))
    n = 1 << 22
    a = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-20, 20, n)).astype(np.float32)
    a[: n // 2] = rng.uniform(-1e5, 1e5, n // 2).astype(np.float32)
    ta = torch.from_numpy(a).cuda()
    tb = torch.full_like(ta, c)
    ieee = H.math_check(1, ta, tb).cpu().numpy()
    q = H.math_check(2, ta, tb).cpu().numpy()
    assert np.array_equal(_bits(q), _bits(ieee))


def test_exp_kernel_equals_library_expf(dev):
    torch, H = dev
    rng = np.random.default_rng(5)
    x = rng.uniform(-80, 80, 1 << 22).astype(np.float32)
    x[:1 << 20] = rng.uniform(-25, 10, 1 << 20).astype(np.float32)   # the arguments used
    tx = torch.from_numpy(x).cuda()
    lib = H.math_check(4, tx).cpu().numpy()
    mine = H.math_check(3, tx).cpu().numpy()
    mism = np.count_nonzero(_bits(mine) != _bits(lib))
    print(f'exp kernel: {mism} of {x.size} differ from expf', flush=True)
    assert mism == 0
    nan = H.math_check(3, torch.tensor([float('nan')], device='cuda')).cpu().numpy()
    assert np.isnan(nan).all()


@pytest.mark.parametrize('y,lo,hi', [(287.04 / 1005.7, 0.005, 1.2),      # Exner, p/p00
                                     (0.2854 * (1.0 - 0.28 * 0.03), 0.8, 25.0),  # theta-e
                                     (0.2854, 0.8, 25.0)])
def test_pow_kernel_accuracy(dev, y, lo, hi):
    torch, H = dev
    rng = np.random.default_rng(3)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), 1 << 22)).astype(np.float32)
    yf = np.float32(y)
    tx = torch.from_numpy(x).cuda()
    ty = torch.full_like(tx, float(yf))
    exact = np.power(x.astype(np.float64), np.float64(yf))
    mine = H.math_check(5, tx, ty).cpu().numpy()
    lib = H.math_check(6, tx, ty).cpu().numpy()
    e_mine, e_lib = _ulp_err(mine, exact), _ulp_err(lib, exact)
    cr = exact.astype(np.float32)
    print(f'pow y={y:.5f}: kernel max {e_mine.max():.3f} ulp, correctly rounded '
          f'{np.mean(mine == cr):.5f}; powf max {e_lib.max():.3f} ulp, correctly rounded '
          f'{np.mean(lib == cr):.5f}', flush=True)
    assert e_mine.max() <= 1.0
    assert e_mine.max() <= max(e_lib.max(), 0.75)


def _ulp_err64(got, x):
    """|got - ln x| in ulps of the double nearest to ln x (long double)."""
    exact = np.log(x.astype(np.longdouble))
    ulp = np.spacing(np.abs(exact.astype(np.float64))).astype(np.longdouble)
    return np.asarray(np.abs(got.astype(np.longdouble) - exact) / ulp, np.float64)


def test_log_kernel_f64_within_an_ulp(dev):
    torch, H = dev
    rng = np.random.default_rng(17)
    n = 1 << 21
    x = 10.0 ** rng.uniform(-307, 308, n)
    q = n // 8
    x[:q] = rng.uniform(0.85, 1.0, q)                   # level pressure ratios
    x[q:2 * q] = rng.uniform(1.0, 1.2, q)
    x[2 * q:3 * q] = rng.uniform(150.0, 350.0, q)      # dewpoints [K]
    # the reduction's edges: around sqrt(1/2) and 1 at every exponent
    e = rng.integers(-1000, 1000, q).astype(np.int32)
    edge = rng.choice([np.sqrt(0.5), 1.0, 2.0 * np.sqrt(0.5)], q)
    x[3 * q:4 * q] = np.ldexp(edge, e) * (1.0 + rng.integers(-4, 5, q) * 2.0 ** -52)
    x[4 * q:4 * q + 1000] = 2.0 ** -1074 * rng.integers(1, 1 << 40, 1000)   # subnormal
    tx = torch.from_numpy(x).cuda()
    mine = H.math_check_f64(0, tx).cpu().numpy()
    lib = H.math_check_f64(1, tx).cpu().numpy()
    e_mine, e_lib = _ulp_err64(mine, x), _ulp_err64(lib, x)
    cr = np.log(x.astype(np.longdouble)).astype(np.float64)
    print(f'log f64 kernel: max {e_mine.max():.3f} ulp, correctly rounded '
          f'{np.mean(mine == cr):.5f}; library max {e_lib.max():.3f} ulp, correctly rounded '
          f'{np.mean(lib == cr):.5f}', flush=True)
    assert e_mine.max() < 1.0
    sp = np.array([0.0, -0.0, np.inf, -np.inf, -1.0, np.nan, 1.0, 2.0])
    got = H.math_check_f64(0, torch.from_numpy(sp).cuda()).cpu().numpy()
    want = np.array([-np.inf, -np.inf, np.inf, np.nan, np.nan, np.nan, 0.0, np.log(2.0)])
    np.testing.assert_array_equal(got, want)


def test_exp_kernel_f64_within_an_ulp(dev):
    torch, H = dev
    rng = np.random.default_rng(19)
    n = 1 << 21
    x = rng.uniform(-708.0, 709.0, n)
    q = n // 4
    x[:q] = rng.uniform(-30.0, 0.0, q)            # the vapour pressure's exponent
    x[q:2 * q] = rng.uniform(-1.0, 1.0, q)
    # around the reduction's half-integer multiples of ln 2
    j = rng.integers(-1000, 1000, q)
    x[2 * q:3 * q] = (j + 0.5) * np.log(2.0) * (1.0 + rng.integers(-8, 9, q) * 2.0 ** -52)
    tx = torch.from_numpy(x).cuda()
    mine = H.math_check_f64(2, tx).cpu().numpy()
    lib = H.math_check_f64(3, tx).cpu().numpy()
    exact = np.exp(x.astype(np.longdouble))
    ulp = np.spacing(exact.astype(np.float64)).astype(np.longdouble)
    e_mine = np.asarray(np.abs(mine.astype(np.longdouble) - exact) / ulp, np.float64)
    e_lib = np.asarray(np.abs(lib.astype(np.longdouble) - exact) / ulp, np.float64)
    cr = exact.astype(np.float64)
    print(f'exp f64 kernel: max {e_mine.max():.3f} ulp, correctly rounded '
          f'{np.mean(mine == cr):.5f}; library max {e_lib.max():.3f} ulp, correctly rounded '
          f'{np.mean(lib == cr):.5f}', flush=True)
    assert e_mine.max() < 1.0
    sp = np.array([np.nan, np.inf, -np.inf, 710.0, -746.0, 0.0, -0.0, -740.0, 709.7])
    got = H.math_check_f64(2, torch.from_numpy(sp).cuda()).cpu().numpy()
    want = H.math_check_f64(3, torch.from_numpy(sp).cuda()).cpu().numpy()
    np.testing.assert_array_equal(got[:7], want[:7])
    np.testing.assert_allclose(got[7:], want[7:], rtol=2.3e-16)
    # subnormal results: within an ulp of the subnormal spacing
    sub = -np.linspace(708.5, 744.0, 1001)
    got = H.math_check_f64(2, torch.from_numpy(sub).cuda()).cpu().numpy()
    ref = np.exp(sub.astype(np.longdouble))
    assert np.all(np.abs(got.astype(np.longdouble) - ref) <= np.longdouble(2.0 ** -1074))


def test_raw_division_f64_is_ieee(dev):
    torch, H = dev
    rng = np.random.default_rng(31)
    n = 1 << 21
    # quotients within 1e-300 .. 1e300 (normal, far from overflow)
    ab = rng.choice([-1.0, 1.0], (n, 2)) * 10.0 ** rng.uniform(-150, 150, (n, 2))
    q = n // 2
    ab[:q, 0] = rng.uniform(-1e5, 1e5, q)           # the heights' operands
    ab[:q, 1] = rng.uniform(0.5, 1.2e5, q)
    raw = H.math_check_f64(4, torch.from_numpy(ab).cuda()).cpu().numpy()
    ieee = H.math_check_f64(5, torch.from_numpy(ab).cuda()).cpu().numpy()
    np.testing.assert_array_equal(ieee, ab[:, 0] / ab[:, 1])
    mism = np.count_nonzero(raw.view(np.int64) != ieee.view(np.int64))
    print(f'raw FP64 division: {mism} of {n} differ from IEEE', flush=True)
    assert mism == 0
    sp = np.array([[1.0, 0.0], [0.0, 0.0], [np.inf, 2.0], [2.0, np.inf], [np.nan, 1.0],
                   [3.0, np.nan]])
    assert np.isnan(H.math_check_f64(4, torch.from_numpy(sp).cuda()).cpu().numpy()).all()

