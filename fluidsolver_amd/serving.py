"""Serving loop over the device library: scene jobs arrive, are admitted
against an HBM budget, advance together in batched launches, and leave.

The reference runs one scene per process (Simulator, src/libFluid/Simulator.cpp:7-57);
there is no serving layer to mirror.  Each job here is exactly what a
Simulator would run for `steps` steps with the scene's ConstantTimestepGenerator
(neighborhood search, then the solver step), and its result is identical to
running the scene alone (fsa_server_* in include/fluidsolver_amd.h).
"""
import ctypes as C

from . import abi
from ._native import lib
from .solver import check


class Server:
    """Continuous batching of scene jobs on one context."""

    def __init__(self, context, hbm_budget_bytes=0, max_running=0, serial_ticks=False):
        cfg = abi.ServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_running = int(max_running)
        cfg.serial_ticks = int(bool(serial_ticks))
        h = C.c_void_p()
        check(lib().fsa_server_create(context.handle, C.byref(cfg), C.byref(h)))
        self.context = context
        self._h = h
        self._collections = {}
        context._servers.add(self)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().fsa_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, desc, collection, steps):
        """Queue a job; the collection's rows are copied, and later receive the
        final state through collect()."""
        jid = C.c_uint64(0)
        v = abi.make_view(collection)
        check(lib().fsa_server_submit(self._h, C.byref(desc), C.byref(v), int(steps), C.byref(jid)))
        self._collections[jid.value] = collection
        return jid.value

    def cancel(self, job_id):
        """Cancels a queued or running job; its state stays readable through job()
        until release().  A running job whose steps have all run is retired DONE
        instead, and its result stays collectable."""
        check(lib().fsa_server_cancel(self._h, int(job_id)))
        if self.job(job_id).state != abi.JOB_DONE:
            self._collections.pop(job_id, None)

    def release(self, job_id):
        """Forgets a job in a terminal state without copying its final state (a failed
        or cancelled job has none; a done job's is dropped)."""
        check(lib().fsa_server_collect(self._h, int(job_id), None))
        self._collections.pop(job_id, None)

    def tick(self):
        """Admit, advance every running scene one step, retire; returns the number of scenes stepped."""
        n = C.c_uint32(0)
        check(lib().fsa_server_tick(self._h, C.byref(n)))
        return n.value

    def job(self, job_id):
        info = abi.JobInfo()
        check(lib().fsa_server_job(self._h, int(job_id), C.byref(info)))
        return info

    def collect(self, job_id, collection=None):
        c = self._collections.pop(job_id) if collection is None else collection
        check(lib().fsa_server_collect(self._h, int(job_id), C.byref(abi.make_view(c))))
        self._collections.pop(job_id, None)
        return c

    def stats(self):
        st = abi.ServerStats()
        check(lib().fsa_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def idle(self):
        st = self.stats()
        return st["jobs_queued"] == 0 and st["jobs_running"] == 0

    def run_until_idle(self):
        """Tick until no job is queued or running.  Every tick either steps a
        job or retires one that cannot be admitted, so this terminates."""
        while not self.idle():
            self.tick()
