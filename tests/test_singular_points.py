"""The compartment formulas near their removable singularities, on the CPU.

beta_r = 0.02 (V + 8.5) / (exp((V + 8.5)/5) - 1) (src/openmp/infoli.c:568) and
alpha_x = 0.13 (V + 25) / (1 - exp(-(V + 25)/10)) (:627, :656) are 0/0 at
V = -8.5 and -25 mV exactly, in the reference as here.  The question is
whether the fast compartment arithmetic (hardware exp2 of x * log2(e), and
a / b as a times the hardware reciprocal of b) can turn a float voltage
*next to* those points into a zero denominator, an infinity or a NaN.  It
cannot: the hardware exp2 and reciprocal are within 1 ulp, so every float V
around each point is evaluated with the correctly rounded exp2 and
reciprocal and with each moved 1 ulp either way (9 combinations), and every
result is finite and positive (within a factor of two of the limit).  Only the exact singular voltage gives NaN,
as the reference's arithmetic does."""
import itertools

import numpy as np
import pytest

f32 = np.float32
LOG2E = f32(1.44269504088896341)


def ulp_moves(x, k):
    x = np.asarray(x, np.float32)
    if k > 0:
        return np.nextafter(x, f32(np.inf))
    if k < 0:
        return np.nextafter(x, f32(-np.inf))
    return x


def exp2_hw(y, k):
    with np.errstate(over="ignore"):
        return ulp_moves(np.exp2(y.astype(np.float64)).astype(np.float32), k)


def rcp_hw(b, k):
    with np.errstate(divide="ignore", over="ignore"):
        return ulp_moves((1.0 / b.astype(np.float64)).astype(np.float32), k)


def D(a, b, k):
    with np.errstate(invalid="ignore", over="ignore"):
        return (a * rcp_hw(b, k)).astype(np.float32)


def E(x, k):
    return exp2_hw((x * LOG2E).astype(np.float32), k)


def neighbourhood(v0, span=1 << 15):
    """Every float within `span` ulps of v0 on either side, v0 included."""
    up, dn = [f32(v0)], [f32(v0)]
    for _ in range(span):
        up.append(np.nextafter(up[-1], f32(np.inf)))
        dn.append(np.nextafter(dn[-1], f32(-np.inf)))
    return np.array(dn[::-1] + up[1:], np.float32)


def beta_r_fm(V, ke, kr1, kr2):
    x = (V + f32(8.5)).astype(np.float32)
    return D((f32(0.02) * x).astype(np.float32), (E(D(x, np.full_like(V, 5), kr1), ke) - f32(1)).astype(np.float32), kr2)


def alpha_x_fm(V, ke, kr1, kr2):
    x = (V + f32(25)).astype(np.float32)
    return D((f32(0.13) * x).astype(np.float32),
             (f32(1) - E(D(-x, np.full_like(V, 10), kr1), ke)).astype(np.float32), kr2)


@pytest.mark.parametrize("name,v0,fn", [("beta_r", -8.5, beta_r_fm), ("alpha_x", -25.0, alpha_x_fm)])
def test_fast_math_finite_next_to_singular_points(name, v0, fn):
    V = neighbourhood(v0)
    at = V == f32(v0)
    for ke, kr1, kr2 in itertools.product((-1, 0, 1), repeat=3):
        out = fn(V, ke, kr1, kr2)
        assert np.all(np.isfinite(out[~at])), (name, ke, kr1, kr2)
        assert np.all(out[~at] > 0), (name, ke, kr1, kr2)
        # the limit at the point is beta_r -> 0.1, alpha_x -> 1.3; next to it
        # exp(x) - 1 cancels (as with expf), which bounds the value within a
        # factor of two of the limit, never 0 or infinite
        lim = 0.1 if name == "beta_r" else 1.3
        assert out[~at].min() > lim / 2 and out[~at].max() < 2 * lim
        if ke == kr1 == kr2 == 0:   # exp2(0) = 1 and rcp(0) = inf are exact on the hardware
            assert np.isnan(out[at]).all()


@pytest.mark.parametrize("v0,which", [(-8.5, "beta_r"), (-25.0, "alpha_x")])
def test_accurate_arithmetic_next_to_singular_points(v0, which):
    """The reference's float arithmetic (expf, IEEE division) behaves the same."""
    V = neighbourhood(v0, 1 << 12)
    x = (V + f32(-v0)).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        if which == "beta_r":
            den = (np.exp((x / f32(5)).astype(np.float64)).astype(np.float32) - f32(1)).astype(np.float32)
            out = ((f32(0.02) * x).astype(np.float32) / den).astype(np.float32)
        else:
            den = (f32(1) - np.exp((-x / f32(10)).astype(np.float64)).astype(np.float32)).astype(np.float32)
            out = ((f32(0.13) * x).astype(np.float32) / den).astype(np.float32)
    at = V == f32(v0)
    assert np.all(np.isfinite(out[~at])) and np.isnan(out[at]).all()
