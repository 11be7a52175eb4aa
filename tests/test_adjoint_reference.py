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
