"""The numpy adjoint reference (tests/adjoint_reference.py) against central
differences of its own forward: it is the reference the device adjoint is
compared with, so its formulas are pinned here (CPU only)."""
import numpy as np
import pytest

import adjoint_reference as R


@pytest.mark.parametrize("ie", [True, False])
def test_reference_adjoint_matches_central_differences(ie):
    fib = R.Fibre(h=np.full(19, 0.01), sb=9, se=12, onset=0.005, duration=0.1, amplitude=1200.0,
                  sensors=[0, 6, 10, 19])
    dt = 1e-3
    steps = np.array([[k * dt, dt, dt, dt] for k in range(30)])
    th_true = np.array([120.0, 36.0, 0.3, 3.828])
    th0 = th_true * np.array([1.03, 0.97, 1.02, 1.04])
    target = R.forward(fib, th_true, steps, ie)
    tr = R.forward(fib, th0, steps, ie)
    grad = R.gradient(fib, th0, steps, ie, tr - target)
    loss = lambda th: 0.5 * np.sum((R.forward(fib, th, steps, ie) - target) ** 2)  # noqa: E731
    for q in range(4):
        e = np.zeros(4)
        e[q] = 1e-6 * th0[q]
        fd = (loss(th0 + e) - loss(th0 - e)) / (2 * e[q])
        assert abs(grad[q] - fd) <= 1e-6 * abs(fd), (q, grad[q], fd)


def test_removable_singularity_derivative_exact():
    """d/dx [x / (E(x) - 1)] with E(x) = (1 + x/1024)^1024 (the HH rates alpha_m,
    alpha_n near V = -50 and -65 mV) against exact rational arithmetic, on both
    sides of the switch from the power series to the quotient rule at |x| = 0.05"""
    from fractions import Fraction
    for xf in (1e-9, -3e-7, 1e-5, 2.5e-3, -0.0123, 0.0499, -0.0499, 0.0501, 0.3, -1.7):
        x = Fraction(xf)
        e = (1 + x / 1024) ** 1024
        ed = (1 + x / 1024) ** 1023
        exact = (e - 1 - x * ed) / (e - 1) ** 2
        got = float(R._du(xf))
        # series: truncation after eight terms (3e-14 at |x| = 0.05); quotient
        # rule: its cancellation bound ~2.5e-13 / x^2
        tol = 1e-13 if abs(xf) < 0.05 else 2.5e-13 / xf ** 2
        assert abs(got - float(exact)) <= tol * abs(float(exact)), (xf, got, float(exact))
    assert float(R._du(0.0)) == -R._C[2]
