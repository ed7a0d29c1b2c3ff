"""Continuous-batching minimisation server on one GPU.

Requests (molecules to minimise, with their FF_Bonded term lists when they
carry them) arrive over time. Each tick the server
  1. admits waiting requests, first come first served, while the resident
     set stays inside the device-memory budget (pdgpu_instance_bytes);
  2. advances every resident minimisation by `steps_per_tick` steepest-
     descent iterations in shared launches (pdgpu_sd_advance); by default
     an instance's list is rebuilt when one of its atoms has moved more
     than half the list padding (update_nlist=-1; PD's every-step rebuild
     is update_nlist=1);
  3. collects the minimisations that finished (force RMS below the
     tolerance, step limit, or failure), frees their slots and memory.
Coordinates, neighbour lists and minimiser state stay on the device for the
whole life of a request; only finished structures come back. Each request
keeps its own step counter and list schedule, so its result is the same
whether it ran alone or among others.
"""
import collections
import time

import numpy as np

from . import Context, J_PER_KCALMOL, PDGPU_ERR_NOMEM, PdgpuError


class Request:
    __slots__ = ("rid", "mol", "arrival", "admitted", "finished", "steps", "energies", "xyz", "iid", "failed")

    def __init__(self, rid, mol, arrival):
        self.rid, self.mol, self.arrival = rid, mol, arrival
        self.admitted = self.finished = None
        self.steps = 0
        self.energies = None
        self.xyz = None
        self.iid = None
        self.failed = False

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class MinimisationServer:
    def __init__(self, options, memory_budget, max_steps=1000, grms_tol=0.1, step_size=2e5, update_nlist=-1,
                 steps_per_tick=4, keep_xyz=True):
        self.ctx = Context(options)
        self.budget = int(memory_budget)
        self.max_steps, self.grms_tol, self.step_size = max_steps, grms_tol, step_size
        self.update_nlist, self.steps_per_tick = update_nlist, steps_per_tick
        self.keep_xyz = keep_xyz
        self.resident = {}          # iid -> Request
        self.resident_bytes = 0
        self.ticks = 0

    def close(self):
        self.ctx.close()

    def _bytes(self, mol):
        b = self.ctx.instance_bytes(len(mol["charge"]), int(mol["mid"]), len(mol.get("sp_i", [])))
        return b + (Context.bonded_bytes(mol) if "bd_idx" in mol else 0)

    def try_admit(self, req, now):
        need = self._bytes(req.mol)
        if need > self.budget:
            raise PdgpuError(PDGPU_ERR_NOMEM, f"request {req.rid} needs {need} B, budget {self.budget} B")
        if self.resident_bytes + need > self.budget:
            return False
        iid = self.ctx.add_system(req.mol, bonded="bd_idx" in req.mol)
        self.ctx.sd_start(iid, self.max_steps, self.step_size, self.update_nlist, self.grms_tol)
        req.iid, req.admitted = iid, now
        self.resident[iid] = req
        self.resident_bytes += need
        return True

    def tick(self, now_fn):
        """Advance all resident minimisations; returns the finished requests."""
        if not self.resident:
            return []
        done, steps, running = self.ctx.sd_advance(self.steps_per_tick)
        self.ticks += 1
        out = []
        for iid in [i for i in self.resident if done[i]]:
            req = self.resident.pop(iid)
            e, st = self.ctx.sd_finish(iid)
            req.energies, req.steps = e, st
            req.failed = st < 0 or e["status"] != 0
            if self.keep_xyz:
                req.xyz = self.ctx.get_positions(iid)
            self.ctx.remove(iid)
            self.resident_bytes -= self._bytes(req.mol)
            req.finished = now_fn()
            out.append(req)
        return out

    def run(self, requests, clock=time.perf_counter):
        """Serves requests (list of Request with arrival times in seconds from
        the start) in real time; returns them with finish times filled in."""
        t0 = clock()
        now = lambda: clock() - t0
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        finished = []
        while pending or waiting or self.resident:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            while waiting and self.try_admit(waiting[0], t):
                waiting.popleft()
            if self.resident:
                finished += self.tick(now)
            elif pending:
                time.sleep(max(0.0, min(0.001, pending[0].arrival - now())))
        return finished


def poisson_arrivals(rate, n, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def summarize(reqs, wall):
    lat = np.array([r.latency for r in reqs if not r.failed])
    return {"requests": len(reqs), "failed": int(sum(r.failed for r in reqs)),
            "minimisations_per_s": len(reqs) / wall, "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_steps": float(np.mean([r.steps for r in reqs])) if reqs else 0.0,
            "final_epot_kcal_mean": float(np.mean([r.energies["epot"] / J_PER_KCALMOL for r in reqs if not r.failed]))
            if len(lat) else None}
