"""The Rust shim's delegation contract, transliterated for tests (test helper).

The shim that binds the C ABI into sherd-core's `Executor` answers every call whose status is not
SHERD_OK from the CPU executor and counts it by kind, the three arms of
crates/sherd-gpu/src/executor.rs:425-487: a batch refused by policy (SHERD_E_LIMIT, SHERD_E_ARG) is
a fall-through counted as delegated; a refused readback (SHERD_E_CORRUPT, a receipt that did not
come back) is counted as corrupt; a device error (SHERD_E_HIP, SHERD_E_NOMEM, SHERD_E_NODEV) as a
host error. The device and the CPU executor are any two objects with the methods below (the
binding's HipExecutor and the oracle, or a fake device on the CPU)."""

import numpy as np

E_ARG, E_HIP, E_NOMEM, E_LIMIT, E_NODEV, E_CORRUPT = -1, -2, -3, -4, -5, -6


class ShimStats:
    def __init__(self):
        self.calls = 0
        self.on_device = 0
        self.delegated = 0
        self.host_errors = 0
        self.corrupt = 0


class Shim:
    def __init__(self, device, cpu, error_type):
        self.device = device
        self.cpu = cpu
        self.error_type = error_type  # the binding's exception carrying `.status`
        self.stats = ShimStats()

    def _delegate(self, status):
        if status == E_CORRUPT:
            self.stats.corrupt += 1
        elif status in (E_HIP, E_NOMEM, E_NODEV):
            self.stats.host_errors += 1
        else:
            self.stats.delegated += 1
        return self.cpu

    def _call(self, method, *args):
        self.stats.calls += 1
        try:
            out = getattr(self.device, method)(*args)
        except self.error_type as e:
            return getattr(self._delegate(e.status), method)(*args)
        self.stats.on_device += 1
        return out

    def coarse_scores(self, *args):
        return self._call("coarse_scores", *args)

    def icp_rung(self, *args):
        return self._call("icp_rung", *args)

    def nearest_below(self, *args):
        return self._call("nearest_below", *args)


class OracleExecutor:
    """The CPU side: the oracle's restatements with the binding's signatures."""

    def __init__(self, O):
        self.O = O

    def coarse_scores(self, tp, tn, pp, pn, poses, radius, normal_agree=0.7):
        return self.O.coarse_scores(tp, tn, pp, pn, poses, radius, normal_agree)

    def icp_rung(self, src, tp, tn, inits, estimation, distance, max_iteration):
        return self.O.icp_rung(src, tp, tn, inits, estimation, distance, max_iteration)

    def nearest_below(self, cloud, q, bound):
        return self.O.nearest_below(cloud, q, bound)


def sheet_case(n=400, k=6, seed=3):
    """A small coarse-score case: a warped sheet with normals, its own points as the probe, and k
    poses near the identity."""
    rng = np.random.default_rng(seed)
    u, v = rng.random(n), rng.random(n)
    P = np.stack([u, v, 0.1 * np.sin(4 * u) * np.cos(3 * v)], 1)
    N = np.stack([-0.4 * np.cos(4 * u) * np.cos(3 * v), 0.3 * np.sin(4 * u) * np.sin(3 * v), np.ones(n)], 1)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    poses = np.repeat(np.eye(4)[None], k, axis=0)
    poses[:, :3, 3] = rng.normal(scale=0.01, size=(k, 3))
    return P, N, poses
