"""The shim's failure contract on the CPU, against a fault-injecting fake device: every status other
than SHERD_OK is answered by the CPU executor (here the oracle) and counted by kind, as
crates/sherd-gpu/src/executor.rs:425-487 counts its three arms; a call that succeeds is the
device's answer. The same contract against the real device's injected faults is tests/
test_gpu_faults.py."""

import numpy as np
import pytest

import oracle_lib as O
import shim


class FakeError(RuntimeError):
    def __init__(self, status):
        super().__init__(f"status {status}")
        self.status = status


class FakeDevice:
    """Answers from the oracle, marked so the test can tell who answered, or fails with the next
    scripted status."""

    def __init__(self, script):
        self.script = list(script)

    def _next(self):
        s = self.script.pop(0) if self.script else 0
        if s:
            raise FakeError(s)

    def coarse_scores(self, *a):
        self._next()
        scores, counts = O.coarse_scores(*a)
        return scores + 1000.0, counts  # the device's mark

    def nearest_below(self, *a):
        self._next()
        idx, dist = O.nearest_below(*a)
        return idx, dist + 1000.0


@pytest.mark.parametrize("status,counter", [(shim.E_HIP, "host_errors"), (shim.E_NOMEM, "host_errors"),
                                            (shim.E_NODEV, "host_errors"), (shim.E_CORRUPT, "corrupt"),
                                            (shim.E_LIMIT, "delegated"), (shim.E_ARG, "delegated")])
def test_each_status_is_answered_by_the_cpu_and_counted(status, counter):
    P, N, poses = shim.sheet_case()
    s = shim.Shim(FakeDevice([0, status, 0]), shim.OracleExecutor(O), FakeError)
    want, wc = O.coarse_scores(P, N, P, N, poses, 0.05)
    a, _ = s.coarse_scores(P, N, P, N, poses, 0.05)   # the device
    b, bc = s.coarse_scores(P, N, P, N, poses, 0.05)  # the scripted failure: the CPU's answer
    c, _ = s.coarse_scores(P, N, P, N, poses, 0.05)   # the device again
    assert np.array_equal(a, want + 1000.0) and np.array_equal(c, want + 1000.0)
    assert np.array_equal(b, want) and np.array_equal(bc, wc)
    st = s.stats
    assert (st.calls, st.on_device) == (3, 2)
    counts = dict(delegated=st.delegated, host_errors=st.host_errors, corrupt=st.corrupt)
    assert counts[counter] == 1 and sum(counts.values()) == 1


def test_a_run_of_failures_is_counted_once_each():
    rng = np.random.default_rng(5)
    cloud = rng.random((300, 3))
    q = rng.random((50, 3))
    script = [shim.E_HIP, shim.E_CORRUPT, 0, shim.E_LIMIT, shim.E_CORRUPT, 0]
    s = shim.Shim(FakeDevice(script), shim.OracleExecutor(O), FakeError)
    want_i, want_d = O.nearest_below(cloud, q, 0.1)
    for k, st in enumerate(script):
        idx, dist = s.nearest_below(cloud, q, 0.1)
        assert np.array_equal(idx, want_i)
        assert np.array_equal(dist, want_d if st else want_d + 1000.0)
    assert (s.stats.host_errors, s.stats.corrupt, s.stats.delegated, s.stats.on_device) == (1, 2, 1, 2)


def test_the_self_test_reduction_is_the_references():
    """selftest.rs:405-460 over 1e7 terms and 256 workgroups gives 0x49a7230c (README.md:1232); the
    same terms in another grouping round otherwise, so the check is a check of the order."""
    assert O.reduce_bits(10_000_000, 256) == 0x49A7230C
    assert O.reduce_bits(10_000_000, 255) != 0x49A7230C
