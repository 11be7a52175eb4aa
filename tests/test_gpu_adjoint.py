"""Discrete adjoint on the GPU: gradients of a Vm-trace misfit with respect to
per-fibre (g_Na, g_K, g_L, sigma), checked against central finite differences
of the same forward (h = 1e-6 relative), through the torch.autograd.Function.
"""
import numpy as np
import pytest

from cases import AM, CM, uneven

pytestmark = pytest.mark.gpu


def _problem(n_fibers=4, n_steps=600, interval=100, scheme="CN", seed=6, lo=200, hi=400, n_sensors=8, tape="auto"):
    from opendihu_amd.adjoint import AdjointProblem
    from opendihu_amd.fibers import FiberBatch
    mesh, stim = uneven(n_fibers, lo=lo, hi=hi, seed=seed)
    n = mesh["n_nodes"]
    mid = n // 2
    stim["stim_begin"] = (mid - 1).astype(np.int32)
    stim["stim_end"] = (mid + 2).astype(np.int32)
    stim["onset"] = np.random.default_rng(seed).uniform(0.0, 0.3, n_fibers)
    b = FiberBatch(**mesh, scheme=scheme, kernels="fused", thomas="partitioned")
    b.set_current_stimulation(**stim)
    dt = 1e-3
    steps = np.stack([np.arange(n_steps) * dt, np.full(n_steps, dt), np.full(n_steps, dt / 2), np.full(n_steps, dt)], 1)
    sensors = np.stack([np.round(np.linspace(0, k - 1, n_sensors)).astype(np.int32) for k in n])
    prob = AdjointProblem(b, steps, sensors, checkpoint_interval=interval, tape=tape)
    rng = np.random.default_rng(seed + 1)
    theta_true = np.stack([rng.uniform(100, 140, n_fibers), rng.uniform(30, 42, n_fibers),
                           rng.uniform(0.2, 0.4, n_fibers), rng.uniform(3.0, 4.5, n_fibers)], 1)
    theta0 = theta_true * rng.uniform(0.95, 1.05, theta_true.shape)
    return prob, theta_true, theta0


def _loss(prob, theta, target):
    tr = prob.forward(theta)
    return 0.5 * float(np.sum((tr - target) ** 2)), tr


@pytest.mark.parametrize("scheme", ["CN", "IE"])
def test_adjoint_matches_central_differences(scheme):
    prob, theta_true, theta0 = _problem(scheme=scheme)
    target = prob.forward(theta_true)
    L0, tr = _loss(prob, theta0, target)
    grad = prob.backward(tr - target)
    assert L0 > 0 and np.all(np.isfinite(grad))
    rng = np.random.default_rng(3)
    v = rng.standard_normal(theta0.shape) * theta0
    h = 1e-6
    lp, _ = _loss(prob, theta0 + h * v, target)
    lm, _ = _loss(prob, theta0 - h * v, target)
    fd = (lp - lm) / (2 * h)
    ad = float(np.sum(grad * v))
    print(f"{scheme}: directional derivative adjoint {ad:.10e} FD {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    # single components: central differences at h (and 2h, Richardson, printed);
    # a small component's difference quotient carries the rounding noise of the
    # whole loss (~1e-5 of the conductivity block here), so each is checked
    # against the largest gradient of its parameter
    def fd(f, q, hq):
        e = np.zeros_like(theta0)
        e[f, q] = hq * theta0[f, q]
        lp, _ = _loss(prob, theta0 + e, target)
        lm, _ = _loss(prob, theta0 - e, target)
        return (lp - lm) / (2 * e[f, q])
    for f, q in [(0, 0), (1, 1), (2, 2), (3, 3), (0, 3)]:
        d1, d2 = fd(f, q, h), fd(f, q, 2 * h)
        dr = (4 * d1 - d2) / 3
        scale = np.max(np.abs(grad[:, q]))
        print(f"  fibre {f} param {q}: adjoint {grad[f, q]:.10e} FD(h) {d1:.10e} FD(2h) {d2:.10e} "
              f"Richardson {dr:.10e}  err/block max {abs(grad[f, q] - d1) / scale:.2e}")
        assert abs(grad[f, q] - d1) <= 1e-5 * scale


@pytest.mark.parametrize("scheme", ["CN", "IE"])
def test_adjoint_exact_on_a_short_horizon(scheme):
    """80 steps, target offset by 1 mV: the adjoint against Richardson-extrapolated
    central differences (h = 1e-5, 2e-5) in random directions. The differences'
    own error here is ~1e-7 (the dense numpy adjoint of the same problem shows
    the same gap to its differences), so the bar is 1e-6; the exact check is
    test_adjoint_equals_dense_reference."""
    prob, theta_true, theta0 = _problem(scheme=scheme, n_steps=80, interval=30)
    target = prob.forward(theta_true) + 1.0
    L0, tr = _loss(prob, theta0, target)
    grad = prob.backward(tr - target)
    rng = np.random.default_rng(11)
    for trial in range(3):
        v = rng.standard_normal(theta0.shape) * theta0
        ad = float(np.sum(grad * v))
        d = []
        for hq in (1e-5, 2e-5):
            lp, _ = _loss(prob, theta0 + hq * v, target)
            lm, _ = _loss(prob, theta0 - hq * v, target)
            d.append((lp - lm) / (2 * hq))
        dr = (4 * d[0] - d[1]) / 3
        print(f"{scheme} direction {trial}: adjoint {ad:.12e} Richardson FD {dr:.12e} rel {abs(ad - dr) / abs(dr):.2e}")
        assert abs(ad - dr) <= 1e-6 * abs(dr)


def test_adjoint_traces_match_forward_run():
    prob, theta_true, _ = _problem(n_steps=300)
    tr = prob.forward(theta_true)
    b = prob.batch
    for k in range(b.n_states):
        b.set_state(k, prob.initial[k])
    b.advance_schedule(prob.steps)
    v = b.vm()
    off = np.concatenate([[0], np.cumsum(b.n_nodes)[:-1]])
    np.testing.assert_array_equal(tr[-1], v[off[:, None] + prob.sensor_nodes])


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from opendihu_amd.adjoint import MonodomainTraces
    prob, theta_true, theta0 = _problem(n_steps=300, interval=64)
    target = torch.from_numpy(prob.forward(theta_true))
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = 0.5 * ((MonodomainTraces.apply(th, prob) - target) ** 2).sum()
    loss.backward()
    tr = prob.forward(theta0)
    expect = prob.backward(tr - target.numpy())
    np.testing.assert_allclose(th.grad.numpy(), expect, rtol=1e-12, atol=0)


@pytest.mark.parametrize("tape", ["whole", "segments"])
def test_checkpoint_interval_does_not_change_the_gradient(tape):
    """bit-identical gradients for one 1300-step segment and for segments of
    37 and 650 steps"""
    prob, theta_true, theta0 = _problem(n_steps=1300, interval=1300, tape=tape)
    target = prob.forward(theta_true)
    tr = prob.forward(theta0)
    assert prob.whole_tape() == (tape == "whole")
    g1 = prob.backward(tr - target)
    for interval in (37, 650):
        prob.interval = interval
        tr = prob.forward(theta0)
        g2 = prob.backward(tr - target)
        np.testing.assert_array_equal(g1, g2, err_msg=f"interval {interval}")


@pytest.mark.parametrize("scheme", ["CN", "IE"])
def test_adjoint_equals_dense_reference(scheme):
    """two fibres (30 and 41 nodes), 60 steps, pulses at different onsets: the
    device gradient against the dense numpy forward + adjoint of the same step
    (which is pinned against its own central differences on the CPU)"""
    import adjoint_reference as R
    from opendihu_amd.adjoint import AdjointProblem
    from opendihu_amd.fibers import FiberBatch
    n = np.array([30, 41], np.int32)
    h = [np.full(k - 1, 0.01) for k in n]
    sb, se = np.array([13, 3], np.int32), np.array([16, 6], np.int32)
    onset = np.array([0.004, 0.011])
    sensors = np.array([[0, 7, 15, 29], [2, 4, 20, 40]], np.int32)
    dt = 1e-3
    steps = np.array([[k * dt, dt, dt, dt] for k in range(60)])
    rng = np.random.default_rng(5)
    theta = np.stack([rng.uniform(100, 140, 2), rng.uniform(30, 42, 2), rng.uniform(0.2, 0.4, 2),
                      rng.uniform(3.0, 4.5, 2)], 1)
    b = FiberBatch(n_nodes=n, element_lengths=np.concatenate(h), prefactor=theta[:, 3] / (AM * CM), scheme=scheme,
                   kernels="fused", thomas="partitioned")
    b.set_current_stimulation(stim_begin=sb, stim_end=se, onset=onset, duration=np.full(2, 0.1),
                              amplitude=np.full(2, 1200.0))
    prob = AdjointProblem(b, steps, sensors, checkpoint_interval=16)
    tr = prob.forward(theta)
    dtr = np.random.default_rng(6).standard_normal(tr.shape)
    grad = prob.backward(dtr)
    for f in range(2):
        fib = R.Fibre(h=h[f], sb=int(sb[f]), se=int(se[f]), onset=onset[f], duration=0.1, amplitude=1200.0,
                      sensors=sensors[f])
        rt = R.forward(fib, theta[f], steps, scheme == "IE", AM * CM)
        np.testing.assert_allclose(tr[:, f, :], rt, rtol=1e-11, atol=1e-10)
        rg = R.gradient(fib, theta[f], steps, scheme == "IE", dtr[:, f, :], AM * CM)
        print(f"{scheme} fibre {f}: device {grad[f]} reference {rg} rel {np.abs(grad[f] - rg) / np.abs(rg)}")
        np.testing.assert_allclose(grad[f], rg, rtol=1e-9)


def test_torch_autograd_on_device_equals_host_path():
    torch = pytest.importorskip("torch")
    from opendihu_amd.adjoint import MonodomainTraces
    prob, theta_true, theta0 = _problem(n_steps=300, interval=64)
    dev = torch.device("cuda")
    target = prob.forward_tensor(torch.tensor(theta_true, device=dev))
    th = torch.tensor(theta0, dtype=torch.float64, device=dev, requires_grad=True)
    tr = MonodomainTraces.apply(th, prob)
    assert tr.is_cuda
    loss = 0.5 * ((tr - target) ** 2).sum()
    loss.backward()
    tr_host = prob.forward(theta0)
    np.testing.assert_array_equal(tr.detach().cpu().numpy(), tr_host)
    expect = prob.backward(tr_host - target.cpu().numpy())
    np.testing.assert_allclose(th.grad.cpu().numpy(), expect, rtol=1e-12, atol=0)


def test_two_outputs_in_one_loss():
    """two MonodomainTraces outputs of one solver in one loss: each output's
    gradient comes from its own forward (the solver holds the checkpoints of the
    last forward only; the earlier output's backward recomputes its forward)"""
    torch = pytest.importorskip("torch")
    from opendihu_amd.adjoint import MonodomainTraces
    prob, theta_true, theta0 = _problem(n_steps=300, interval=64)
    dev = torch.device("cuda")
    target = prob.forward_tensor(torch.tensor(theta_true, device=dev))
    theta1 = theta0 * np.array([1.03, 0.98, 1.01, 0.97])
    ta = torch.tensor(theta0, dtype=torch.float64, device=dev, requires_grad=True)
    tb = torch.tensor(theta1, dtype=torch.float64, device=dev, requires_grad=True)
    loss = 0.5 * ((MonodomainTraces.apply(ta, prob) - target) ** 2).sum() \
        + 0.25 * ((MonodomainTraces.apply(tb, prob) - target) ** 2).sum()
    loss.backward()
    ga = prob.backward(prob.forward(theta0) - target.cpu().numpy())
    gb = prob.backward(0.5 * (prob.forward(theta1) - target.cpu().numpy()))
    np.testing.assert_allclose(ta.grad.cpu().numpy(), ga, rtol=1e-12, atol=0)
    np.testing.assert_allclose(tb.grad.cpu().numpy(), gb, rtol=1e-12, atol=0)


def test_tape_mode_errors_are_reported():
    """asking for the tape mode before any forward, a backward before any
    forward and an unknown tape mode are errors, not crashes"""
    prob, theta_true, theta0 = _problem(n_steps=50, interval=16)
    with pytest.raises(RuntimeError, match="fms_adjoint_forward first"):
        prob.whole_tape()
    with pytest.raises(RuntimeError, match="fms_adjoint_forward first"):
        prob.backward(np.zeros((50, 4, 8)))
    with pytest.raises(RuntimeError, match="adjoint tape"):
        prob.batch.set_tuning(adjoint_tape=2)
    with pytest.raises(ValueError):
        prob.set_tape("disk")
    prob.forward(theta0)
    assert prob.whole_tape()


@pytest.mark.parametrize("scheme", ["CN", "IE"])
def test_whole_tape_equals_checkpoints(scheme):
    """the forward recording the whole tape (the backward only reverses it)
    and the forward keeping checkpoints (the backward recomputes every segment,
    the last one partial): the same traces and bit-identical gradients; at this
    size the automatic choice is the whole tape"""
    prob, theta_true, theta0 = _problem(n_steps=500, interval=64, scheme=scheme, tape="auto")
    target = prob.forward(theta_true)
    assert prob.whole_tape()
    out = {}
    # with the whole tape the forward and the reverse sweep run in launches of
    # the solver's steps per fused launch (512: one launch; 48: ten and a
    # partial one), with checkpoints in the 64-step segments
    for tape, per_launch in (("whole", 512), ("segments", 512), ("whole", 48)):
        prob.set_tape(tape)
        prob.batch.set_tuning(steps_per_launch=per_launch)
        tr = prob.forward(theta0)
        assert prob.whole_tape() == (tape == "whole")
        g = prob.backward(tr - target)
        if tape in out:
            np.testing.assert_array_equal(tr, out[tape][0])
            np.testing.assert_array_equal(g, out[tape][1])
        out[tape] = (tr, g)
    prob.batch.set_tuning(steps_per_launch=512)
    np.testing.assert_array_equal(out["whole"][0], out["segments"][0])
    np.testing.assert_array_equal(out["whole"][1], out["segments"][1])
    assert np.all(np.isfinite(out["whole"][1])) and np.any(out["whole"][1] != 0)
    choice = prob.tune_backward(theta0, out["whole"][0] - target)
    assert choice["tape"] == "whole" and choice["schedule"] == "whole_tape" and len(choice["seconds"]) == 2


def test_adjoint_schedules_give_the_same_gradient():
    """the segment recompute beside the reverse sweep (with or without raised
    priority) or before it: bit-identical gradients"""
    from opendihu_amd import _lib as L
    prob, theta_true, theta0 = _problem(n_steps=500, interval=64, tape="segments")
    target = prob.forward(theta_true)
    tr = prob.forward(theta0)
    grads = []
    for mode in (L.FMS_ADJOINT_OVERLAP, L.FMS_ADJOINT_OVERLAP_PRIORITY, L.FMS_ADJOINT_SERIAL,
                 L.FMS_ADJOINT_OVERLAP_RECOMPUTE_PRIORITY):
        prob.batch.set_tuning(adjoint=mode)
        grads.append(prob.backward(tr - target))
    for g in grads[1:]:
        np.testing.assert_array_equal(g, grads[0])
    choice = prob.tune_backward(theta0, tr - target)
    assert choice["schedule"] in ("overlap", "overlap_priority", "serial", "overlap_recompute_priority")
    assert choice["threads"] in (256, 512) and len(choice["seconds"]) == 8 and choice["tape"] == "segments"
    # the reverse sweep with 256 or 512 threads per fibre: the partitioned solve
    # takes other chunks and the prefactor gradient is summed over another
    # number of waves, so the gradients agree to rounding
    prob.batch.set_tuning(adjoint_threads=256)
    g256 = prob.backward(tr - target)
    prob.batch.set_tuning(adjoint_threads=512)
    np.testing.assert_allclose(g256, grads[0], rtol=1e-9)
