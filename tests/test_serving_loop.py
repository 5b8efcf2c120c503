"""Host logic of the request loop (simbody_amd.serving.serve) on the CPU,
against a stand-in server that finishes each request after a known number
of advances: first-come-first-served admission limited by free slots,
per-request results routed back to the right request, latency measured from
arrival, and every slot returned."""
import numpy as np

from simbody_amd.serving import serve


class FakeRequests:
    def __init__(self, n, work):
        self.n = n
        self.q = np.arange(n, dtype=np.float64)[:, None] * np.ones((1, 2))
        self.u = -self.q[:, :1].copy()
        self.work = np.asarray(work)          # advances each request needs


class FakeServer:
    def __init__(self, slots):
        self.slots, self.nq, self.nu = slots, 2, 1
        self.free = list(range(slots))[::-1]
        self.busy = {}                        # slot -> [request, advances left]
        self.admitted_order = []
        self.max_busy = 0

    def admit_requests(self, req, i0, k):
        k = min(k, len(self.free))
        out = []
        for i in range(i0, i0 + k):
            s = self.free.pop()
            self.busy[s] = [i, int(req.work[i])]
            self.admitted_order.append(i)
            out.append(s)
        self.max_busy = max(self.max_busy, len(self.busy))
        self.req = req
        return np.array(out, np.int32)

    def advance(self, steps):
        for v in self.busy.values():
            v[1] -= 1

    def collect(self):
        done = [s for s, v in self.busy.items() if v[1] <= 0]
        r = {"slots": np.array(done, np.int32)}
        ids = [self.busy[s][0] for s in done]
        r["q"] = self.req.q[ids]; r["u"] = self.req.u[ids]
        r["status"] = np.ones(len(done), np.int32)
        r["stats"] = np.tile(np.array([7, 6, 1, 0], np.int32), (len(done), 1))
        r["t"] = np.zeros(len(done))
        for s in done:
            del self.busy[s]
            self.free.append(s)
        return r


def test_fifo_admission_capacity_and_routing():
    rng = np.random.default_rng(3)
    n = 300
    work = rng.integers(1, 6, size=n)
    req = FakeRequests(n, work)
    sv = FakeServer(slots=16)
    arrivals = np.cumsum(rng.exponential(1e-4, size=n))
    r = serve(sv, req, arrivals, steps_per_advance=1, time_limit=30)
    assert r["completed"] == n
    assert sv.admitted_order == list(range(n))          # first come, first served
    assert sv.max_busy <= 16 and r["max_occupied"] <= 16
    assert np.array_equal(r["q"], req.q) and np.array_equal(r["u"], req.u)   # results routed to their request
    assert (r["status"] == 1).all() and (r["stats"][:, 1] == 6).all()
    assert np.isfinite(r["latency"]).all() and (r["latency"] >= 0).all()
    assert (r["admitted_at"] >= arrivals - 1e-9).all()   # never admitted before arriving
    assert len(sv.free) == 16


def test_idle_server_waits_for_arrivals():
    req = FakeRequests(3, [1, 1, 1])
    sv = FakeServer(slots=4)
    arrivals = np.array([0.0, 0.05, 0.10])
    r = serve(sv, req, arrivals, steps_per_advance=1, time_limit=10)
    assert r["completed"] == 3
    assert (r["admitted_at"] >= arrivals - 1e-9).all()
    assert r["wall"] >= 0.10
