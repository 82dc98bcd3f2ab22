"""The self-test gate and the failure contract on the device (GPU): sherd_selftest's four checks
(crates/sherd-gpu/src/selftest.rs:179) with the reference's reduction bits, and every fault
sherd_fault_inject can fire: the entry point returns its status, the library counts it, the
context keeps working (the next call's answer is the clean one, bit for bit), and the shim answers
the failed call from the CPU (the oracle) and counts it (executor.rs:425-487)."""

import numpy as np
import pytest

import oracle_lib as O
import shim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


def test_selftest_passes_with_the_references_reduction(ex):
    say("case: sherd_selftest")
    r = ex.selftest()
    for name, ok, detail in r["checks"]:
        say(f"  {name}: {'ok' if ok else 'FAILED'} - {detail}")
    assert r["passed"] and [c[0] for c in r["checks"]] == ["limits", "reduction", "bounded nn", "throughput"]
    assert r["reduce_device_bits"] == r["reduce_host_bits"] == 0x49A7230C
    assert r["queries"] == 6000 * 64 and r["gpu_ms"] > 0 and r["cpu_ms"] > 0


def _coarse(ex):
    P, N, poses = shim.sheet_case()
    return ex.coarse_scores(P, N, P, N, poses, 0.05)


@pytest.mark.parametrize("kind,status,counter", [("host", -2, "host_errors"), ("nomem", -3, "nomem"),
                                                 ("corrupt", -6, "corrupt")])
def test_each_fault_fails_one_call_and_the_context_recovers(ex, kind, status, counter):
    import sherd_hip

    say(f"case: injected {kind} fault on sherd_coarse_scores")
    clean, clean_c = _coarse(ex)
    h0 = ex.health()
    ex.fault_inject(kind)
    with pytest.raises(sherd_hip.HipError) as err:
        _coarse(ex)
    assert err.value.status == status
    again, again_c = _coarse(ex)
    assert np.array_equal(again, clean) and np.array_equal(again_c, clean_c)
    h1 = ex.health()
    d = {k: h1[k] - h0[k] for k in h1}
    say(f"  {err.value}; counters moved {d}")
    assert d["calls"] == 2 and d["failed"] == 1 and d[counter] == 1 and d["injected"] == 1


def test_a_fault_fires_at_its_countdown(ex):
    import sherd_hip

    say("case: host fault armed two calls ahead")
    clean, _ = _coarse(ex)
    ex.fault_inject("host", after=2)
    assert np.array_equal(_coarse(ex)[0], clean) and np.array_equal(_coarse(ex)[0], clean)
    with pytest.raises(sherd_hip.HipError):
        _coarse(ex)
    assert np.array_equal(_coarse(ex)[0], clean)
    ex.fault_inject("corrupt")
    ex.fault_inject("none")  # disarmed: nothing fires
    assert np.array_equal(_coarse(ex)[0], clean)


def test_corrupt_receipts_on_the_batch_entry_points(ex):
    """The ICP rung and the bounded-NN search refuse their readback when the receipt is stale and
    answer cleanly on the next call."""
    import sherd_hip

    rng = np.random.default_rng(9)
    P, N, poses = shim.sheet_case(n=800, k=4)
    inits = poses.reshape(-1, 16)
    say("case: corrupt receipt on sherd_icp_rung and sherd_nearest_below")
    want = ex.icp_rung(P, P, N, inits, "point_to_plane", 0.05, 10)
    ex.fault_inject("corrupt")
    with pytest.raises(sherd_hip.HipError) as err:
        ex.icp_rung(P, P, N, inits, "point_to_plane", 0.05, 10)
    assert err.value.status == -6
    assert np.array_equal(ex.icp_rung(P, P, N, inits, "point_to_plane", 0.05, 10), want)
    q = rng.random((500, 3))
    wi, wd = ex.nearest_below(P, q, 0.1)
    ex.fault_inject("corrupt")
    with pytest.raises(sherd_hip.HipError):
        ex.nearest_below(P, q, 0.1)
    gi, gd = ex.nearest_below(P, q, 0.1)
    assert np.array_equal(gi, wi) and np.array_equal(gd, wd)


def test_the_shim_answers_injected_faults_from_the_cpu(ex):
    import sherd_hip

    say("case: the shim over the device with injected faults")
    P, N, poses = shim.sheet_case()
    want, want_c = O.coarse_scores(P, N, P, N, poses, 0.05)
    s = shim.Shim(ex, shim.OracleExecutor(O), sherd_hip.HipError)
    for kind in ("corrupt", "host", "nomem"):
        ex.fault_inject(kind)
        got, got_c = s.coarse_scores(P, N, P, N, poses, 0.05)  # the CPU's answer
        assert np.array_equal(got, want) and np.array_equal(got_c, want_c)
        dev, dev_c = s.coarse_scores(P, N, P, N, poses, 0.05)  # the device's
        assert np.array_equal(dev_c, want_c) and np.array_equal(dev, want)
    st = s.stats
    assert (st.calls, st.on_device, st.corrupt, st.host_errors, st.delegated) == (6, 3, 1, 2, 0)
