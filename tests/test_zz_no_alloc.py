"""fms_advance_schedule_noalloc (the step table and flags in the library's fixed
device scratch array) against fms_advance_schedule on the same inputs: two calls
in a row, the second with fewer steps, bit-identical states, spikes and
stimulation log; a call larger than the scratch is refused with an error."""
import numpy as np
import pytest

from cases import uneven, vc_lane_stim_range, firing_table, c2

pytestmark = pytest.mark.gpu


def _table(t0, dt, n):
    return np.ascontiguousarray(np.stack([t0 + np.arange(n) * dt, np.full(n, dt), np.full(n, dt / 2),
                                          np.full(n, dt)], 1))


def _firing(batch, mesh):
    n = mesh["n_nodes"]
    F = n.size
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    sb, se = vc_lane_stim_range(off, n // 2, n, 4)
    fe = firing_table(40, 10, 4.0, 0.1, seed=2)
    mu = (np.arange(F) * 7) % 10 + 1
    batch.set_firing_table_stimulation(sb, se, fe, mu, call_frequency=1.0, repeat_after_first_call=0.1,
                                       call_enable_begin=0.2 * (np.arange(F) % 3), jitter_lists=[[0.1, -0.1]] * F,
                                       log_capacity=4096)


def _advance(b, steps, noalloc):
    from opendihu_amd import _lib as L
    fn = b.lib.fms_advance_schedule_noalloc if noalloc else b.lib.fms_advance_schedule
    L.check(b.lib, fn(b.handle, L.as_ptr(steps, L._dp), steps.shape[0], 1), "advance")


@pytest.mark.parametrize("kernels,thomas", [("split", "serial"), ("split", "partitioned"), ("fused", "partitioned")])
def test_noalloc_advance_equals_advance(kernels, thomas):
    from opendihu_amd.fibers import FiberBatch
    mesh, _ = uneven(12, lo=200, hi=600, seed=9)
    runs = []
    for noalloc in (False, True):
        b = FiberBatch(**mesh, scheme="IE", kernels=kernels, thomas=thomas)
        _firing(b, mesh)
        b.enable_spike_detection(0.0)
        dt = 2e-3
        out = []
        for t0, n in ((0.0, 700), (700 * dt, 260)):
            _advance(b, _table(t0, dt, n), noalloc)
            out.append(b.get_states())
        runs.append((out, b.spikes(), b.stimulation_log()))
        b.close()
    for call, (a, c) in enumerate(zip(runs[0][0], runs[1][0])):
        np.testing.assert_array_equal(a, c, err_msg=f"call {call}")
    for a, c in zip(runs[0][1], runs[1][1]):
        np.testing.assert_array_equal(a, c)
    la = sorted(zip(runs[0][2][1].tolist(), runs[0][2][0].tolist()))
    lb = sorted(zip(runs[1][2][1].tolist(), runs[1][2][0].tolist()))
    assert la == lb and len(la) > 0


def test_noalloc_advance_large_model_and_limit():
    """the 56-state model (per-fibre reaction, paired halves), then a call whose
    table does not fit the scratch"""
    import os
    from opendihu_amd import _lib as L
    from opendihu_amd.build import LIB_DIR, SHORTEN_LIBRARY
    from opendihu_amd.fibers import FiberBatch
    if not os.path.exists(os.path.join(LIB_DIR, SHORTEN_LIBRARY)):
        pytest.skip("Shorten library not built")
    mesh, stim = c2(n_fibers=6, n_nodes=301, length_cm=3.0, dt=2e-4, end_time=0.2, rate_hz=4000.0)
    runs = []
    for noalloc in (False, True):
        b = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="partitioned", library=SHORTEN_LIBRARY)
        b.set_firing_table_stimulation(**stim)
        out = []
        for t0, n in ((0.0, 300), (300 * 2e-4, 100)):
            _advance(b, _table(t0, 2e-4, n), noalloc)
            out.append(b.get_states())
        runs.append(out)
        if noalloc:
            too_many = _table(0.0, 2e-4, 200_000)
            rc = b.lib.fms_advance_schedule_noalloc(b.handle, L.as_ptr(too_many, L._dp), too_many.shape[0], 1)
            assert rc == L.FMS_ERR_UNSUPPORTED and "scratch" in L.last_error(b.lib)
            np.testing.assert_array_equal(b.get_states(), out[-1])
        b.close()
    for call, (a, c) in enumerate(zip(*runs)):
        np.testing.assert_array_equal(a, c, err_msg=f"call {call}")
