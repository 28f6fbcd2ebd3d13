"""The serving loop over the device library's C ABI (svxg_server_*, include/svxg.h).

A Server takes solve requests for StressProblems at any time (submit), from any thread; its two
native workers assemble what arrives and step one running batch into which requests are admitted,
and from which finished ones leave, every few PCG steps (continuous batching). Completed requests
come back through poll() / wait() with their results and timestamps on the server's clock.
"""
import ctypes
import threading

import numpy as np

from .abi import BREAK, SVXG_MIXED, ServeDone, ServerOptions, ServerStats, ptr
from .stress import SvxgError, _check, default_context, lib


class Server:
    def __init__(self, ctx=None, precision=SVXG_MIXED, max_active=4096, rtol=1e-8, maxit=5000,
                 amg_min_nodes=160, window=8, restart_iters=120, rebuild_iters=60, setup_workers=2,
                 solve_restarts=60, restart_diverging=True, judge=None):
        """judge: None (requests are not judged) or judge options (stress.default_judge_options):
        each converged request is judged on its solution (the counters in its record, the picks
        through picks())."""
        self.ctx = ctx or default_context()
        o = ServerOptions()
        lib().svxg_server_default_options(ctypes.byref(o))
        o.precision, o.max_active, o.rtol, o.maxit = precision, max_active, rtol, maxit
        o.amg_min_nodes, o.window = amg_min_nodes, window
        o.restart_iters, o.rebuild_iters, o.setup_workers = restart_iters, rebuild_iters, setup_workers
        o.solve_restarts, o.restart_diverging = solve_restarts, int(bool(restart_diverging))
        if judge is not None:
            o.judge = 1
            o.judge_opt = judge
        h = ctypes.c_void_p()
        _check(lib().svxg_server_create(self.ctx.h, ctypes.byref(o), ctypes.byref(h)), "svxg_server_create")
        self.h = h
        self._mu = threading.Lock()
        self._inflight = {}  # ticket -> (problem, u array or None)
        self._done = {}
        self._buf = (ServeDone * 4096)()

    def now_ms(self):
        return lib().svxg_server_now_ms(self.h)

    def submit(self, problem, f=None, warm=False, remove=None, want_u=True):
        """Queue one solve of `problem` (its bonds `remove` leave K first). Returns the ticket."""
        n = problem.n
        fa = None if f is None else np.ascontiguousarray(f, np.float64).reshape(-1)
        if fa is not None and fa.size != 6 * n:
            raise ValueError("f: 6 values per node")
        u = np.zeros(6 * n) if want_u else None
        rm = None if remove is None or len(remove) == 0 else np.ascontiguousarray(remove, np.int32)
        t = ctypes.c_uint64()
        with self._mu:
            _check(lib().svxg_server_submit(self.h, problem.h, ptr(fa), ptr(u), int(bool(warm)), ptr(rm),
                                            0 if rm is None else int(rm.size), ctypes.byref(t)), "server submit")
            self._inflight[t.value] = (problem, u)
        return t.value

    def submit_many(self, problems, fs=None, warm=None, removes=None, want_u=False):
        """Queue len(problems) solves in one call (fs[i] None: the resident load). Returns tickets."""
        n = len(problems)
        if n == 0:
            return []
        fa = [None if fs is None or fs[i] is None else np.ascontiguousarray(fs[i], np.float64) for i in range(n)]
        us = [np.zeros(6 * p.n) if want_u else None for p in problems]
        rms = [None if removes is None or removes[i] is None or len(removes[i]) == 0
               else np.ascontiguousarray(removes[i], np.int32) for i in range(n)]
        vp = ctypes.c_void_p
        ph = (vp * n)(*[p.h.value for p in problems])
        fp = (vp * n)(*[None if a is None else a.ctypes.data for a in fa])
        up = (vp * n)(*[None if a is None else a.ctypes.data for a in us])
        wa = np.array([0] * n if warm is None else [int(bool(w)) for w in warm], np.int32)
        rp = (vp * n)(*[None if a is None else a.ctypes.data for a in rms])
        nr = np.array([0 if a is None else a.size for a in rms], np.int32)
        tk = np.zeros(n, np.uint64)
        with self._mu:
            _check(lib().svxg_server_submit_many(self.h, n, ph, fp, up, ptr(wa), rp, ptr(nr), ptr(tk)),
                   "server submit_many")
            for i in range(n):
                self._inflight[int(tk[i])] = (problems[i], us[i])
        return [int(t) for t in tk]

    def poll(self, wait_ms=0, cap=4096):
        """Completed requests since the last poll (up to cap, at most 4096): {ticket: record}."""
        cap = min(int(cap), 4096)
        buf = self._buf
        n = ctypes.c_int32()
        _check(lib().svxg_server_poll(self.h, buf, cap, ctypes.byref(n), int(wait_ms)), "server poll")
        out = {}
        with self._mu:
            for k in range(n.value):
                d = buf[k]
                problem, u = self._inflight.pop(d.ticket)
                out[d.ticket] = {
                    "problem": problem, "u": u, "status": d.status, "assembled": bool(d.assembled),
                    "result": d.result.as_dict(), "t_submit_ms": d.t_submit_ms, "t_ready_ms": d.t_ready_ms,
                    "t_admit_ms": d.t_admit_ms, "t_done_ms": d.t_done_ms, "restarts": d.restarts,
                    "abandoned": bool(d.abandoned), "breakdowns": d.breakdowns, "detached": d.detached,
                    "judge": d.judge.as_dict(),
                }
        return out

    def wait(self, tickets, timeout_ms=120000):
        """The records of these tickets (others completed meanwhile are kept for later calls)."""
        want = set(tickets)
        got = {t: self._done.pop(t) for t in list(want) if t in self._done}
        t0 = self.now_ms()
        while len(got) < len(want):
            for t, rec in self.poll(wait_ms=50).items():
                (got if t in want else self._done)[t] = rec
            if self.now_ms() - t0 > timeout_ms:
                raise SvxgError(f"server: {len(want) - len(got)} requests still in flight after {timeout_ms} ms")
        return got

    def picks(self, problem, cap=4096):
        """The picks of problem's last judgement by this server (BREAK records, in order)."""
        out = np.zeros(cap, BREAK)
        n = ctypes.c_int32()
        _check(lib().svxg_server_picks(self.h, problem.h, out.ctypes.data, cap, ctypes.byref(n)), "server picks")
        return out[:min(cap, n.value)].copy()

    def release(self, problem):
        """Destroy `problem` once its request (if any) completes."""
        _check(lib().svxg_server_release(self.h, problem.h), "server release")
        problem.h = None

    def stats(self):
        s = ServerStats()
        _check(lib().svxg_server_stats_get(self.h, ctypes.byref(s)), "server stats")
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().svxg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
