"""Dense numpy restatement of the split-step forward and its discrete adjoint
for Hodgkin-Huxley fibres with current-pulse stimulation (test infrastructure).

It follows the same step as the library (fast_monodomain_solver_compute.tpp:
110-126: Heun half at t, one tridiagonal diffusion solve A V^b = B V^a with the
row assembly of compute.tpp:336-429, Heun half at t + dt/2) with dense
matrices and numpy.linalg.solve, and reverses it with the formulas

  Heun:      lambda = lambda' + dt/2 mu + dt/2 J(y)^T (lambda' + dt mu),  mu = J(y*)^T lambda'
             theta_bar += dt/2 [f_theta(y)^T (lambda' + dt mu) + f_theta(y*)^T lambda']
  diffusion: w = A^-1 lambda_b,  lambda_a = B w,  pf_bar += w^T (dB/dpf V^a - dA/dpf V^b)

Its gradients agree with central differences of its own forward to ~1e-8
(h = 1e-6), so it is a reference for the device adjoint on small problems.
"""
from __future__ import annotations

import numpy as np

ENA, EK, EL = -75.0 + 115.0, -75.0 - 12.0, -75.0 + 10.613


def _E(x):
    y = 1.0 + x / 1024.0
    for _ in range(10):
        y = y * y
    return y


def _Ed(x):
    return _E(x) / (1.0 + x / 1024.0)


# coefficients of E(x) - 1 = sum_k C[k] x^k, C[k] = binom(1024, k) / 1024^k
_C = [0.0, 1.0]
for _k in range(1, 8):
    _C.append(_C[-1] * (1024 - _k) / ((_k + 1) * 1024))


def _du(x):
    """d/dx [x / (E(x) - 1)]: the quotient rule, except for |x| < 0.05, where its
    numerator E - 1 - x E' (about -x^2/2) cancels catastrophically and both it
    and E - 1 are summed from their power series (eight terms)"""
    x = np.asarray(x, float)
    with np.errstate(divide="ignore", invalid="ignore"):
        den = _E(x) - 1
        direct = (den - x * _Ed(x)) / den ** 2
        sden = sum(_C[k] * x ** k for k in range(1, 9))
        snum = -sum((k - 1) * _C[k] * x ** k for k in range(2, 9))
        series = np.where(x == 0, -_C[2], snum / sden ** 2)
    return np.where(np.abs(x) < 0.05, series, direct)


def rhs(y, p):
    V, m, h, n = y
    ist, gna, gk, gl = p
    am = (-0.1 * (V + 50)) / (_E(-(V + 50) / 10) - 1)
    bm = 4 * _E(-(V + 75) / 18)
    ah = 0.07 * _E(-(V + 75) / 20)
    bh = 1 / (_E(-(V + 45) / 10) + 1)
    an = (-0.01 * (V + 65)) / (_E(-(V + 65) / 10) - 1)
    bn = 0.125 * _E((V + 75) / 80)
    return np.array([-(-ist + gna * m ** 3 * h * (V - ENA) + gk * n ** 4 * (V - EK) + gl * (V - EL)),
                     am * (1 - m) - bm * m, ah * (1 - h) - bh * h, an * (1 - n) - bn * n])


def vjp(y, p, w):
    """J(y)^T w and f_theta(y)^T w for theta = (gNa, gK, gL)"""
    V, m, h, n = y
    _, gna, gk, gl = p
    xm = -(V + 50) / 10
    num, den = -0.1 * (V + 50), _E(xm) - 1
    am, dam = num / den, -0.1 * _du(xm)
    xbm = -(V + 75) / 18
    bm, dbm = 4 * _E(xbm), 4 * _Ed(xbm) * (-1 / 18)
    xah = -(V + 75) / 20
    ah, dah = 0.07 * _E(xah), 0.07 * _Ed(xah) * (-1 / 20)
    xbh = -(V + 45) / 10
    ebh = _E(xbh)
    bh, dbh = 1 / (ebh + 1), -(_Ed(xbh) * (-0.1)) / (ebh + 1) ** 2
    xn = -(V + 65) / 10
    numn, denn = -0.01 * (V + 65), _E(xn) - 1
    an, dan = numn / denn, -0.01 * _du(xn)
    xbn = (V + 75) / 80
    bn, dbn = 0.125 * _E(xbn), 0.125 * _Ed(xbn) / 80
    out = np.array([
        w[0] * (-(gna * m ** 3 * h + gk * n ** 4 + gl)) + w[1] * (dam * (1 - m) - dbm * m)
        + w[2] * (dah * (1 - h) - dbh * h) + w[3] * (dan * (1 - n) - dbn * n),
        w[0] * (-3 * gna * m ** 2 * h * (V - ENA)) + w[1] * (-am - bm),
        w[0] * (-gna * m ** 3 * (V - ENA)) + w[2] * (-ah - bh),
        w[0] * (-4 * gk * n ** 3 * (V - EK)) + w[3] * (-an - bn)])
    gp = np.array([w[0] * (-m ** 3 * h * (V - ENA)), w[0] * (-n ** 4 * (V - EK)), w[0] * (-(V - EL))])
    return out, gp


def heun(y, p, dt):
    r = rhs(y, p)
    return y + 0.5 * dt * (r + rhs(y + dt * r, p))


def heun_reverse(y, p, dt, lam):
    ys = y + dt * rhs(y, p)
    mu, gs = vjp(ys, p, lam)
    jy, g3 = vjp(y, p, lam + dt * mu)
    return lam + 0.5 * dt * mu + 0.5 * dt * jy, 0.5 * dt * (g3 + gs)


def matrices(h, pf, dt, ie):
    """dense A, B of the split step's diffusion system (compute.tpp:336-429)"""
    n = len(h) + 1
    A, B = np.zeros((n, n)), np.zeros((n, n))
    for i in range(n):
        if i > 0:
            hl = h[i - 1]
            kl, ml, kr, mr = pf / hl, hl / 6, -pf / hl, hl / 3
            if ie:
                A[i, i - 1] = kl - ml / dt; A[i, i] += kr - mr / dt; B[i, i - 1] = -ml / dt; B[i, i] += -mr / dt
            else:
                A[i, i - 1] = kl / 2 - ml / dt; A[i, i] += kr / 2 - mr / dt
                B[i, i - 1] = -kl / 2 - ml / dt; B[i, i] += -kr / 2 - mr / dt
        if i < n - 1:
            hr = h[i]
            kr, mr, kl, ml = pf / hr, hr / 6, -pf / hr, hr / 3
            if ie:
                A[i, i + 1] = kr - mr / dt; A[i, i] += kl - ml / dt; B[i, i + 1] = -mr / dt; B[i, i] += -ml / dt
            else:
                A[i, i + 1] = kr / 2 - mr / dt; A[i, i] += kl / 2 - ml / dt
                B[i, i + 1] = -kr / 2 - mr / dt; B[i, i] += -kl / 2 - ml / dt
    return A, B


def _k1(v, h):
    out = np.zeros(len(v))
    out[1:] += (v[:-1] - v[1:]) / h
    out[:-1] += (v[1:] - v[:-1]) / h
    return out


class Fibre:
    """one fibre: element lengths h, stimulated nodes [sb, se) with a pulse of
    `amplitude` on [onset, onset + duration), sensor nodes"""

    def __init__(self, h, sb, se, onset, duration, amplitude, sensors, y0=(-75.0, 0.05, 0.6, 0.325)):
        self.h = np.asarray(h, float)
        self.n = len(self.h) + 1
        self.sb, self.se, self.onset, self.duration, self.amplitude = sb, se, onset, duration, amplitude
        self.sensors = np.asarray(sensors)
        self.y0 = np.asarray(y0, float)

    def p(self, theta, i, t):
        on = (t >= self.onset - 1e-13 and t < self.onset + self.duration - 1e-13) and self.sb <= i < self.se
        return (self.amplitude if on else 0.0, theta[0], theta[1], theta[2])


def forward(fib: Fibre, theta, steps, ie, am_cm=500.0, tape=None):
    """theta = (gNa, gK, gL, sigma); steps rows (t, dt_split, dt0D, dt1D); one Heun step per half"""
    Y = np.tile(fib.y0[:, None], (1, fib.n))
    traces = []
    for t, dts, dt0, dt1 in steps:
        A, B = matrices(fib.h, theta[3] / am_cm, dt1, ie)
        y0 = Y.copy()
        for i in range(fib.n):
            Y[:, i] = heun(Y[:, i], fib.p(theta, i, t), dt0)
        va = Y[0].copy()
        Y[0] = np.linalg.solve(A, B @ Y[0])
        if tape is not None:
            tape.append((y0, va, Y.copy()))
        tm = t + 0.5 * dts
        for i in range(fib.n):
            Y[:, i] = heun(Y[:, i], fib.p(theta, i, tm), dt0)
        traces.append(Y[0, fib.sensors].copy())
    return np.array(traces)


def gradient(fib: Fibre, theta, steps, ie, dtraces, am_cm=500.0):
    """dL/d(gNa, gK, gL, sigma) for dL/dtraces"""
    tape = []
    forward(fib, theta, steps, ie, am_cm, tape)
    lam = np.zeros((4, fib.n))
    g = np.zeros(3)
    pf_bar = 0.0
    for s in range(len(steps) - 1, -1, -1):
        t, dts, dt0, dt1 = steps[s]
        y0, va, yb = tape[s]
        A, B = matrices(fib.h, theta[3] / am_cm, dt1, ie)
        lam[0, fib.sensors] += dtraces[s]
        tm = t + 0.5 * dts
        for i in range(fib.n):
            lam[:, i], gg = heun_reverse(yb[:, i], fib.p(theta, i, tm), dt0, lam[:, i])
            g += gg
        w = np.linalg.solve(A, lam[0])
        pf_bar += w @ (-_k1(yb[0], fib.h)) if ie else w @ (-0.5 * _k1(va + yb[0], fib.h))
        lam[0] = B @ w
        for i in range(fib.n):
            lam[:, i], gg = heun_reverse(y0[:, i], fib.p(theta, i, t), dt0, lam[:, i])
            g += gg
    return np.concatenate([g, [pf_bar / am_cm]])
