"""Accuracy of the library's elementary functions (evaluated on the host from the same
source the kernels compile): error bounds against float64 numpy, special values, and
the domain of every call site of the transport. The device evaluates the same source
bit-for-bit (GPU tier)."""
import numpy as np
import pytest

import oracle_py


def ulp_err(got, ref):
    """|got - ref| in units of the float ulp at ref (ref in float64)."""
    got = got.astype(np.float64)
    a = np.abs(ref)
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -126)))
    ulp = np.exp2(e - 23)
    return np.abs(got - ref) / ulp


@pytest.fixture(scope="module")
def evald():
    rng = np.random.default_rng(11)
    n = 400_000
    x = np.concatenate([rng.uniform(-20, 20, n // 4), rng.uniform(0, 2 * np.pi, n // 4),
                        np.exp(rng.uniform(-40, 4, n // 4)), rng.uniform(-8000, 8000, n - 3 * (n // 4))])
    y = rng.uniform(-1.0, 1.0, n)
    od, of = oracle_py.fmath_eval(x, y)
    xf = x.astype(np.float32).astype(np.float64)
    yf = y.astype(np.float32).astype(np.float64)
    return x, y, xf, yf, od, of


def test_float_functions_within_2p5_ulp(evald):
    x, y, xf, yf, od, of = evald
    pos = np.abs(xf) + np.float64(np.float32(1e-30))
    assert ulp_err(of[0], np.log(pos.astype(np.float32).astype(np.float64))).max() <= 2.5
    m = (xf > -100) & (xf < 88)
    assert ulp_err(of[1][m], np.exp(xf[m])).max() <= 2.5
    # powf at its call site: taumin^u with taumin in [1/3, 1], u in (0, 1]
    b = np.minimum(np.abs(yf), 1.0)
    m = (b >= 1 / 3) & (np.abs(xf) <= 1.0)
    assert m.sum() > 10_000
    assert ulp_err(of[2][m], np.power(b[m], np.abs(xf[m]))).max() <= 3.0
    # sin/cos: relative ulp error where the value is not tiny, absolute otherwise
    for k, f in ((3, np.sin), (4, np.cos)):
        ref = f(xf)
        assert ulp_err(of[k], ref).max() <= 2.5
    assert ulp_err(of[5], np.arctan2(xf, yf)).max() <= 2.5
    assert ulp_err(of[7], np.arccos(np.clip(yf, -1, 1))).max() <= 2.5


def test_double_functions_within_4_ulp(evald):
    x, y, xf, yf, od, of = evald
    def ulpd(got, ref):
        return np.abs(got - ref) / np.spacing(np.abs(ref))
    assert ulpd(od[0], np.log(np.abs(x) + 1e-300)).max() <= 4
    m = x < 700
    assert ulpd(od[1][m], np.exp(x[m])).max() <= 4
    m = np.abs(np.sin(x)) > 1e-3
    assert ulpd(od[2][m], np.sin(x[m])).max() <= 4
    m = np.abs(np.cos(x)) > 1e-3
    assert ulpd(od[3][m], np.cos(x[m])).max() <= 4
    assert ulpd(od[4], np.arctan2(x, y)).max() <= 6
    assert ulpd(od[5], np.arccos(np.clip(y, -1, 1))).max() <= 6


def test_special_values():
    x = np.array([0.0, -0.0, -1.0, np.inf, -np.inf, np.nan, 1.0, 2.0 ** -149, 1e-45, 88.7, -103.5, -200.0])
    y = np.array([-1.0, -1.0, 0.0, -np.inf, np.inf, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    od, of = oracle_py.fmath_eval(x, y)
    # logf(|x| + 1e-30), expf(x)
    assert of[1][0] == 1.0 and of[1][3] == np.inf and of[1][4] == 0.0 and np.isnan(of[1][5])
    assert of[1][11] == 0.0
    assert of[1][10] == np.float32(np.exp(-103.5))       # gradual underflow, one rounding
    assert np.isnan(of[0][5]) and of[0][3] == np.inf and of[0][6] == 0.0
    # atan2f(x, y): signed zeros and infinities
    assert of[5][0] == np.float32(np.pi) and of[5][1] == -np.float32(np.pi)
    assert of[5][2] == np.float32(-np.pi / 2)
    assert of[5][3] == np.float32(3 * np.pi / 4) and of[5][4] == np.float32(-np.pi / 4)
    # acosf(clip(y)): exact end points
    assert of[7][0] == np.float32(np.pi) and of[7][6] == np.float32(np.pi / 2) and of[7][7] == 0.0


def test_transport_call_site_domains():
    """Woodcock step -ln(u) for RANECU outputs u = k 2^-31, and the Compton profile
    exponent range: tight errors where the physics lives."""
    k = np.unique(np.concatenate([np.arange(1, 5000), np.random.default_rng(3).integers(1, 2 ** 31 - 86, 200_000)]))
    u = (k.astype(np.float32) * np.float32(2.0 ** -31)).astype(np.float64)
    od, of = oracle_py.fmath_eval(u, np.zeros_like(u))
    assert ulp_err(of[0], np.log(u)).max() <= 1.0
    t = np.linspace(-60, 0.5, 100_001)
    od, of = oracle_py.fmath_eval(t, np.zeros_like(t))
    tf = t.astype(np.float32).astype(np.float64)
    assert ulp_err(of[1], np.exp(tf)).max() <= 1.0
