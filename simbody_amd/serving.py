"""Continuous batching of rollout requests on one GPU.

`Server` wraps the tdyn_server_* C ABI (include/tdyn.h): a fixed set of
device slots, sized from an HBM budget, for trees of one chain topology.
Each request brings its own model constants, initial state and final time;
its State stays on the device from admission until it is collected.
`serve()` is the request loop: requests arriving (by their arrival times)
are admitted whenever slots are free, all occupied slots advance together
by a few integrator steps per launch, and finished ones are retired.
Latency is measured from a request's arrival to the moment its result is
on the host.
"""
import ctypes
import time

import numpy as np

from .matter import IntegratorCfgC, TreeDescC, _check, _d, _i32, _stream, fill_tree_desc, lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_d)


class Server:
    def __init__(self, topology, integ, hbm_budget_bytes, max_slots=0, device="cuda:0"):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        keep = []
        d = fill_tree_desc(TreeDescC(), topology, keep)
        cfg = integ.c()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(lib().tdyn_server_create(ctypes.byref(d), ctypes.byref(cfg), ctypes.c_size_t(int(hbm_budget_bytes)),
                                            int(max_slots), ctypes.byref(h)))
        self.h = h
        self.nb = len(topology.parent)
        inf = self.info()
        self.slots, self.nq, self.nu, self.bytes_per_slot = inf["slots"], inf["nq"], inf["nu"], inf["bytes_per_slot"]
        # host result buffers, allocated once (collect returns views of them)
        m = self.slots
        self._out = dict(slots=np.zeros(m, np.int32), q=np.zeros((m, self.nq)), u=np.zeros((m, self.nu)),
                         t=np.zeros(m), status=np.zeros(m, np.int32), stats=np.zeros((m, 4), np.int32))

    def __del__(self):
        if getattr(self, "h", None) is not None:
            try:
                from . import matter
                if matter._lib is not None:
                    matter._lib.tdyn_server_destroy(self.h)
            except ImportError:      # interpreter shutdown: the process exit frees the device memory
                pass
            self.h = None

    def info(self):
        a = (ctypes.c_int64 * 6)()
        _check(lib().tdyn_server_info(self.h, a))
        return dict(zip(["slots", "free", "occupied", "bytes_per_slot", "nq", "nu"], list(a)))

    def admit(self, specs, q, u, t_end, fapp=None):
        """Admit as many of the requests as there are free slots (in order);
        returns the slots of the admitted ones (len <= len(specs))."""
        n = len(specs)
        if n == 0:
            return np.zeros(0, np.int32)
        keep = []
        descs = (TreeDescC * n)()
        for i, s in enumerate(specs):
            fill_tree_desc(descs[i], s, keep)
        q = np.ascontiguousarray(q, np.float64).reshape(n, self.nq)
        u = np.ascontiguousarray(u, np.float64).reshape(n, self.nu)
        t_end = np.ascontiguousarray(t_end, np.float64).reshape(n)
        fa = None if fapp is None else np.ascontiguousarray(fapp, np.float64).reshape(n, -1)
        slots = np.zeros(n, np.int32)
        k = ctypes.c_int(0)
        _check(lib().tdyn_server_admit(self.h, n, descs, _dp(q), _dp(u), _dp(t_end), _dp(fa),
                                       slots.ctypes.data_as(_i32), ctypes.byref(k), _stream()))
        return slots[:k.value]

    def admit_requests(self, req, i0, k):
        """Admit requests i0..i0+k-1 of a Requests (as many as there are free
        slots); returns their slots."""
        return req.admit_range(self, i0, k)

    def advance(self, max_steps):
        _check(lib().tdyn_server_advance(self.h, int(max_steps), _stream()))

    def collect(self, max_out=None):
        """Retire finished requests.  The returned arrays are views of buffers
        that the next collect() overwrites: copy what must be kept."""
        m = self.slots if max_out is None else min(int(max_out), self.slots)
        o = self._out
        k = ctypes.c_int(0)
        _check(lib().tdyn_server_collect(self.h, m, o["slots"].ctypes.data_as(_i32), _dp(o["q"]), _dp(o["u"]),
                                         _dp(o["t"]), o["status"].ctypes.data_as(_i32),
                                         o["stats"].ctypes.data_as(_i32), ctypes.byref(k), _stream()))
        k = k.value
        return {key: v[:k] for key, v in o.items()}


class TreeServer:
    """Slots for requests of different trees (tdyn_tree_server_*): any size,
    topology and mobilizer mix; each request's memory comes from the
    server's pools, sized from the HBM budget, by its own tree's needs.
    lanes: 0 the level-parallel evaluator where a tree allows it (lanes from
    its widest level), 1 the one-lane tree sweeps, 4..64 fixed lanes."""

    def __init__(self, integ, hbm_budget_bytes, max_slots=0, device="cuda:0", lanes=0):
        import torch
        self.device = torch.device(device)
        cfg = integ.c()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(lib().tdyn_tree_server_create(ctypes.byref(cfg), ctypes.c_size_t(int(hbm_budget_bytes)),
                                                 int(max_slots), ctypes.byref(h)))
        self.h = h
        self.set_lanes(lanes)
        self.slots = self.info()["slots"]
        self._ny = np.zeros(self.slots, np.int64)     # state size of the request in each slot
        self._y = np.zeros(1 << 12)
        self._off = np.zeros(self.slots, np.int64)
        self._sl = np.zeros(self.slots, np.int32)
        self._t = np.zeros(self.slots)
        self._st = np.zeros(self.slots, np.int32)
        self._stats = np.zeros((self.slots, 4), np.int32)

    def __del__(self):
        if getattr(self, "h", None) is not None:
            try:
                from . import matter
                if matter._lib is not None:
                    matter._lib.tdyn_tree_server_destroy(self.h)
            except ImportError:
                pass
            self.h = None

    def info(self):
        a = (ctypes.c_int64 * 7)()
        _check(lib().tdyn_tree_server_info(self.h, a))
        return dict(zip(["slots", "free", "occupied", "budget_bytes", "bytes_in_use", "body_pool", "state_pool"],
                        list(a)))

    def set_lanes(self, lanes):
        """Evaluator of the requests admitted from now on (see the class)."""
        _check(lib().tdyn_tree_server_set_lanes(self.h, int(lanes)))

    @staticmethod
    def request_bytes(spec):
        keep = []
        d = fill_tree_desc(TreeDescC(), spec, keep)
        b = ctypes.c_int64(0)
        _check(lib().tdyn_tree_server_request_bytes(ctypes.byref(d), ctypes.byref(b)))
        return b.value

    def admit(self, specs, q, u, t_end, fapp=None):
        """Admit requests in order while a slot is free and the pools have
        room; q, u, fapp: per-request arrays (lists) or packed.  Returns the
        slots of the admitted ones (the first len(result) requests)."""
        n = len(specs)
        if n == 0:
            return np.zeros(0, np.int32)
        keep = []
        descs = (TreeDescC * n)()
        for i, sp in enumerate(specs):
            fill_tree_desc(descs[i], sp, keep)

        def pk(x):
            if isinstance(x, (list, tuple)):
                x = np.concatenate([np.ravel(v) for v in x])
            return np.ascontiguousarray(x, np.float64).reshape(-1)
        q = pk(q); u = pk(u)
        t_end = np.ascontiguousarray(t_end, np.float64).reshape(n)
        fa = None if fapp is None else pk(fapp)
        slots = np.zeros(n, np.int32)
        k = ctypes.c_int(0)
        _check(lib().tdyn_tree_server_admit(self.h, n, descs, _dp(q), _dp(u), _dp(t_end), _dp(fa),
                                            slots.ctypes.data_as(_i32), ctypes.byref(k), _stream()))
        slots = slots[:k.value]
        for i, sl in enumerate(slots):
            self._ny[sl] = specs[i].nq + specs[i].nu
        return slots

    def admit_requests(self, req, i0, k):
        """Admit requests i0..i0+k-1 of a TreeRequests (in order, while a
        slot is free and the pools have room); returns their slots."""
        return req.admit_range(self, i0, k)

    def advance(self, max_steps):
        _check(lib().tdyn_tree_server_advance(self.h, int(max_steps), _stream()))

    def collect(self, max_out=None):
        """Retire finished requests: slots, t, status, stats and the list of
        their states y (q then u), copies."""
        m = self.slots if max_out is None else min(int(max_out), self.slots)
        need = int(self._ny.sum())
        if self._y.size < need:
            self._y = np.zeros(max(need, 2 * self._y.size))
        k = ctypes.c_int(0)
        _check(lib().tdyn_tree_server_collect(self.h, m, ctypes.c_int64(self._y.size),
                                              self._sl.ctypes.data_as(_i32), _dp(self._y),
                                              self._off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              _dp(self._t), self._st.ctypes.data_as(_i32),
                                              self._stats.ctypes.data_as(_i32), ctypes.byref(k), _stream()))
        kk = k.value
        sl = self._sl[:kk].copy()
        ys = [self._y[self._off[i]:self._off[i] + self._ny[sl[i]]].copy() for i in range(kk)]
        self._ny[sl] = 0
        return {"slots": sl, "t": self._t[:kk].copy(), "status": self._st[:kk].copy(),
                "stats": self._stats[:kk].copy(), "y": ys}


class Requests:
    """n rollout requests with their tree descriptions prebuilt, so that the
    request loop admits a contiguous run of them with one C call and no
    per-request host work."""

    def __init__(self, specs, q, u, t_end, fapp=None):
        n = len(specs)
        self.n = n
        self._keep = []
        self.descs = (TreeDescC * n)()
        for i, s in enumerate(specs):
            fill_tree_desc(self.descs[i], s, self._keep)
        self._init_state(q, u, t_end, fapp)

    @classmethod
    def from_chain_params(cls, template, params, q, u, t_end, fapp=None, model_of=None):
        """Requests whose trees share `template`'s topology and take their
        constants from models.chain_params output (no per-request arrays):
        request i uses row model_of[i] (default i) of the parameters, and
        row i of q, u, t_end."""
        self = cls.__new__(cls)
        mass, com, inert, X_PF, X_BM = (np.ascontiguousarray(a, np.float64) for a in params)
        nm, nb = mass.shape
        n = len(t_end)
        model_of = np.arange(n) if model_of is None else np.asarray(model_of)
        assert len(model_of) == n and (model_of < nm).all()
        self.n = n
        par = np.ascontiguousarray(template.parent, np.int32); mob = np.ascontiguousarray(template.mobilizer, np.int32)
        self._keep = [par, mob, mass, com, inert, X_PF, X_BM]
        self.descs = (TreeDescC * n)()
        addr = [a.ctypes.data for a in (mass, com, inert, X_PF, X_BM)]
        step = [8 * nb, 24 * nb, 48 * nb, 96 * nb, 96 * nb]
        pp, pm = par.ctypes.data_as(_i32), mob.ctypes.data_as(_i32)
        g = [float(x) for x in template.gravity]
        for i in range(n):
            d = self.descs[i]
            d.nb = nb; d.parent = pp; d.mobilizer = pm
            m = int(model_of[i])
            d.mass, d.com_B, d.inertia_B, d.X_PF, d.X_BM = (ctypes.cast(a + m * st, _d) for a, st in zip(addr, step))
            d.gravity[0], d.gravity[1], d.gravity[2] = g
        self._init_state(q, u, t_end, fapp)
        return self

    def _init_state(self, q, u, t_end, fapp):
        n = self.n
        self.q = np.ascontiguousarray(q, np.float64).reshape(n, -1)
        self.u = np.ascontiguousarray(u, np.float64).reshape(n, -1)
        self.t_end = np.ascontiguousarray(t_end, np.float64).reshape(n)
        self.fapp = None if fapp is None else np.ascontiguousarray(fapp, np.float64).reshape(n, -1)

    def admit_range(self, server, i0, k):
        """Admit requests i0..i0+k-1 (as many as there are free slots)."""
        if k <= 0:
            return np.zeros(0, np.int32)
        sz = ctypes.sizeof(TreeDescC)
        descs = ctypes.cast(ctypes.addressof(self.descs) + i0 * sz, ctypes.POINTER(TreeDescC))
        qp = ctypes.cast(self.q.ctypes.data + i0 * self.q.shape[1] * 8, _d)
        up = ctypes.cast(self.u.ctypes.data + i0 * self.u.shape[1] * 8, _d)
        tp = ctypes.cast(self.t_end.ctypes.data + i0 * 8, _d)
        fp = None if self.fapp is None else ctypes.cast(self.fapp.ctypes.data + i0 * self.fapp.shape[1] * 8, _d)
        slots = np.zeros(k, np.int32)
        kk = ctypes.c_int(0)
        _check(lib().tdyn_server_admit(server.h, int(k), descs, qp, up, tp, fp, slots.ctypes.data_as(_i32),
                                       ctypes.byref(kk), _stream()))
        return slots[:kk.value]


class TreeRequests:
    """n requests of different trees for a TreeServer, prebuilt (tree
    descriptions, packed states and forces with per-request offsets) so the
    request loop admits a contiguous run of them with one C call."""

    def __init__(self, specs, q, u, t_end, fapp=None):
        n = len(specs)
        self.n = n
        self.specs = specs
        self._keep = []
        self.descs = (TreeDescC * n)()
        for i, sp in enumerate(specs):
            fill_tree_desc(self.descs[i], sp, self._keep)
        self.ny = np.array([sp.nq + sp.nu for sp in specs], np.int64)
        self.q = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in q]), np.float64)
        self.u = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in u]), np.float64)
        self.qo = np.concatenate([[0], np.cumsum([sp.nq for sp in specs])]).astype(np.int64)
        self.uo = np.concatenate([[0], np.cumsum([sp.nu for sp in specs])]).astype(np.int64)
        self.t_end = np.ascontiguousarray(t_end, np.float64).reshape(n)
        if fapp is None:
            self.fapp = None
        else:
            self.fapp = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in fapp]), np.float64)
            self.fo = np.concatenate([[0], np.cumsum([6 * (sp.nb + 1) for sp in specs])]).astype(np.int64)

    def admit_range(self, server, i0, k):
        if k <= 0:
            return np.zeros(0, np.int32)
        sz = ctypes.sizeof(TreeDescC)
        descs = ctypes.cast(ctypes.addressof(self.descs) + i0 * sz, ctypes.POINTER(TreeDescC))
        qp = ctypes.cast(self.q.ctypes.data + int(self.qo[i0]) * 8, _d)
        up = ctypes.cast(self.u.ctypes.data + int(self.uo[i0]) * 8, _d)
        tp = ctypes.cast(self.t_end.ctypes.data + i0 * 8, _d)
        fp = None if self.fapp is None else ctypes.cast(self.fapp.ctypes.data + int(self.fo[i0]) * 8, _d)
        slots = np.zeros(k, np.int32)
        kk = ctypes.c_int(0)
        _check(lib().tdyn_tree_server_admit(server.h, int(k), descs, qp, up, tp, fp, slots.ctypes.data_as(_i32),
                                            ctypes.byref(kk), _stream()))
        slots = slots[:kk.value]
        server._ny[slots] = self.ny[i0:i0 + kk.value]
        return slots


def serve(server, req, arrivals, steps_per_advance=16, time_limit=None, warmup=0.0):
    """Run the request loop over `req` (a Requests for a Server, or a
    TreeRequests for a TreeServer), whose arrival times (seconds from the
    start, sorted) are `arrivals`.  Requests are admitted first come, first
    served, as slots (and, for a TreeServer, pool memory) allow.  Returns
    per-request results (q and u, or y for trees; status, stats, latency)
    and loop figures; the steady-state figures count the requests that
    arrive after `warmup` seconds and before the last arrival."""
    n = req.n
    trees = isinstance(server, TreeServer)
    arrivals = np.asarray(arrivals, np.float64)
    assert np.all(np.diff(arrivals) >= 0), "arrivals must be sorted"
    if trees:
        out_y = [None] * n
    else:
        out_q = np.zeros((n, server.nq)); out_u = np.zeros((n, server.nu))
    status = np.zeros(n, np.int32); stats = np.zeros((n, 4), np.int32)
    latency = np.full(n, np.nan); admitted_at = np.full(n, np.nan); done_at = np.full(n, np.nan)
    slot_req = np.full(server.slots, -1, np.int64)
    head = 0            # next request to admit (FIFO); requests head..nxt-1 are queued
    nxt = 0; n_done = 0; occupied = 0
    launches = 0; max_occupied = 0
    t_admit = 0.0; t_step = 0.0       # host wall time in admission / in advance + collect
    t0 = time.perf_counter()
    while n_done < n:
        now = time.perf_counter() - t0
        if time_limit is not None and now > time_limit:
            break
        nxt = int(np.searchsorted(arrivals, now, side="right"))
        free = server.slots - occupied
        if nxt > head and free > 0:
            k = min(free, nxt - head)
            ta = time.perf_counter()
            slots = server.admit_requests(req, head, k)
            t_admit += time.perf_counter() - ta
            kk = len(slots)
            slot_req[slots] = np.arange(head, head + kk)
            admitted_at[head:head + kk] = time.perf_counter() - t0
            head += kk
            occupied += kk
            max_occupied = max(max_occupied, occupied)
        if occupied == 0:
            if nxt < n:
                wait = arrivals[min(nxt, n - 1)] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            continue
        ta = time.perf_counter()
        server.advance(steps_per_advance)
        launches += 1
        res = server.collect()
        tdone = time.perf_counter() - t0
        t_step += time.perf_counter() - ta
        ids = slot_req[res["slots"]]
        slot_req[res["slots"]] = -1
        if trees:
            for j, i in enumerate(ids):
                out_y[i] = res["y"][j]
        else:
            out_q[ids] = res["q"]; out_u[ids] = res["u"]
        status[ids] = res["status"]; stats[ids] = res["stats"]
        latency[ids] = tdone - arrivals[ids]
        done_at[ids] = tdone
        k = len(ids)
        n_done += k; occupied -= k
    wall = time.perf_counter() - t0
    out = dict(status=status, stats=stats, latency=latency, admitted_at=admitted_at, done_at=done_at,
               wall=wall, launches=launches, max_occupied=max_occupied, completed=n_done,
               admit_s=t_admit, advance_collect_s=t_step)
    if trees:
        out["y"] = out_y
    else:
        out["q"] = out_q; out["u"] = out_u
    # steady state: requests arriving in [warmup, last arrival]
    t1 = float(arrivals[-1]) if n else 0.0
    win = (arrivals >= warmup) & (arrivals <= t1)
    if t1 > warmup and win.any():
        ok = win & np.isfinite(latency)
        out["steady"] = dict(window_s=t1 - warmup, offered_req_s=float(win.sum()) / (t1 - warmup),
                             completed_in_window_req_s=float(((done_at >= warmup) & (done_at <= t1)).sum())
                             / (t1 - warmup),
                             requests=int(win.sum()), completed=int(ok.sum()),
                             latency_p50_s=float(np.percentile(latency[ok], 50)) if ok.any() else None,
                             latency_p99_s=float(np.percentile(latency[ok], 99)) if ok.any() else None)
    return out
