"""Request loop that serves XC Fock builds to QM/MM molecular-dynamics replicas.

Each replica is one QM region whose geometry and density move every MD step.
A request asks for the XC energy and Fock matrix of one replica at its next
step.  The loop does continuous batching on one GPU:

* every replica owns a context (its own HIP stream and device arena); the
  grid, partition and basis values of its current geometry stay on the
  device between its steps, and are rebuilt in place (host staging, device
  Becke weights, device basis evaluation) when the step moves the atoms;
* requests are admitted as soon as a slot frees up and their replica is not
  already being served; each admitted request runs on a worker thread (its
  host staging, its grid rebuild and its build on the replica's stream), so
  the host work of several requests and the device work of all admitted
  requests overlap;
* admission is bounded by an HBM budget: idle replicas' arenas are released,
  least recently used first, when a new one would not fit; with ``slab`` the
  budget is reserved once as a device slab and every arena is carved from it
  on the host, so a request makes no device allocation;
* latency is arrival -> result on the host, reported as p50/p99 together with
  Fock builds per second.

Results do not depend on what else is running: each build is computed by the
same deterministic kernels on its own buffers, so a replica's step gives the
same numbers alone and in the mix.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import threading
import time

import numpy as np

from . import g2g, systems

ANG = 1.0 / systems.ANG_PER_BOHR  # bohr per angstrom


@dataclasses.dataclass
class Request:
    rid: int
    replica: int
    step: int
    arrival: float
    start: float = 0.0
    done: float = 0.0
    energy: float = 0.0
    electrons: float = 0.0
    fock: np.ndarray | None = None
    set_ms: float = 0.0     # problem arrays handed to the context
    stage_ms: float = 0.0   # host staging of the new grid
    load_ms: float = 0.0    # device rebuild (allocation, uploads, weights, basis values)
    solve_ms: float = 0.0   # the build itself, queued to collected
    load_phases: list | None = None  # weights, planning, uploads, evaluation (ms)


class Replica:
    """One QM region evolving by a seeded random walk: per step the atoms move
    by N(0, 0.01 A) per coordinate and the density by a symmetric N(0, 1e-4)
    update, after which the density is rescaled to its initial trace."""

    def __init__(self, system: systems.System, seed: int, pos_sigma_ang: float = 0.01,
                 rho_sigma: float = 1e-4):
        self.base = system
        self.seed = seed
        self.pos_sigma = pos_sigma_ang * ANG
        self.rho_sigma = rho_sigma
        self.trace0 = float(np.trace(system.P))
        self._step = None  # cached walk position (step, positions, density)
        self._lock = threading.Lock()
        self._ahead = None  # (step, future) of a state computed ahead

    def state(self, step: int) -> systems.System:
        """The system at MD step ``step`` (step 0 is the initial one).  The walk
        is a fixed seeded sequence, so a step is the same whoever asks; asking
        for the steps in order advances it by one step each time."""
        with self._lock:
            return self._state(step)

    def prefetch(self, step: int, pool: concurrent.futures.Executor):
        """Starts computing state(step) on a worker thread (the next MD step
        of a replica is known as soon as its current one is taken)."""
        self._ahead = (step, pool.submit(self.state, step))

    def take(self, step: int) -> systems.System:
        """state(step), from the prefetched computation when there is one."""
        ahead, self._ahead = self._ahead, None
        if ahead is not None and ahead[0] == step:
            return ahead[1].result()
        return self.state(step)

    def _state(self, step: int) -> systems.System:
        if self._step is None or step < self._step:
            self._rng = np.random.default_rng((self.seed, 0))
            self._step, self._pos, self._P = 0, self.base.pos.copy(), self.base.P.copy()
        M = self._P.shape[0]
        while self._step < step:
            self._pos = self._pos + self._rng.normal(0.0, self.pos_sigma, size=self._pos.shape)
            D = self._rng.normal(0.0, self.rho_sigma, size=(M, M))
            P = self._P + 0.5 * (D + D.T)
            self._P = P * (self.trace0 / np.trace(P))
            self._step += 1
        return dataclasses.replace(self.base, pos=self._pos.copy(), P=self._P.copy(),
                                   rho_packed=systems.pack_density(self._P))


class Server:
    """Continuous-batching XC Fock server over a set of replicas (one GPU).

    The main thread admits a request when a slot is free, its replica has no
    build in flight and its arena fits the HBM budget, and hands it to a
    worker thread, which rebuilds the replica's grid for the new geometry
    (host staging, device Becke weights and basis values), queues the build
    on the replica's own stream and waits for it.  The library releases the
    interpreter lock in every call, so the workers' host staging and their
    device builds overlap, and builds of different replicas are in flight
    together on their streams.
    """

    def __init__(self, replicas: list[Replica], precision: str = "bf16x6", slots: int = 8,
                 hbm_budget: int | None = None, grid_mode: str = "device",
                 density_variant: int | None = None, slab: bool = False):
        self.replicas = replicas
        self.precision = precision
        self.density_variant = density_variant
        self.grid_mode = grid_mode
        self.slots = slots
        self.budget = hbm_budget
        self.ctx: list[g2g.Context | None] = [None] * len(replicas)
        self.arena = [0] * len(replicas)        # bytes of each replica's last grid
        self.resident = [False] * len(replicas)
        self.last_use = [0.0] * len(replicas)
        self.busy = [False] * len(replicas)
        self.evictions = 0
        self.retries = 0  # builds that found the device full and were queued again
        self.next_step = [0] * len(replicas)  # MD steps served so far, per replica
        if slab:  # one device slab of the whole budget: no allocation per rebuild
            if hbm_budget is None:
                raise ValueError("a slab needs an HBM budget")
            g2g.reserve_slab(hbm_budget)
        self.workers = concurrent.futures.ThreadPoolExecutor(max_workers=slots)
        # the next MD step's inputs of each replica are computed ahead on
        # other threads (numpy releases the interpreter lock in its loops)
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=4)
        self.last_step = [0] * len(replicas)

    # -- device memory -------------------------------------------------------
    def _estimate(self, r: int) -> int:
        """Arena bytes a replica needs: its last grid's size, or for a replica
        never built a bound from its size (points ~ 7000 per atom, a quarter
        of its functions significant per point, 10 FP64 planes)."""
        if self.arena[r]:
            return int(1.1 * self.arena[r])
        s = self.replicas[r].base
        return int(7000 * s.natom * max(1, s.M // 4) * 10 * 8 * 1.3)

    def _resident_bytes(self) -> int:
        return sum(a for a, res in zip(self.arena, self.resident) if res)

    def _make_room(self, r: int) -> bool:
        """Releases idle replicas (least recently used first) until replica r
        fits the budget; False if it cannot fit now (builds in flight hold the
        memory).  A replica larger than the whole budget runs once nothing
        else holds device memory."""
        if self.budget is None:
            return True
        need = 0 if self.resident[r] else self._estimate(r)
        while self._resident_bytes() + need > self.budget:
            idle = [k for k in range(len(self.replicas))
                    if self.resident[k] and not self.busy[k] and k != r]
            if not idle:
                others = self._resident_bytes() - (self.arena[r] if self.resident[r] else 0)
                return others == 0
            k = min(idle, key=lambda q: self.last_use[q])
            self.ctx[k].release_device()
            self.resident[k] = False
            self.evictions += 1
        return True

    # -- one request ---------------------------------------------------------
    @staticmethod
    def _build(req: Request, ctx: g2g.Context, s: systems.System):
        """Worker thread: rebuild the grid for the step's geometry and run
        the build; returns (arena bytes, energy, F, info), or None when the
        device ran out of memory (the replica then holds no arena)."""
        req.start = time.perf_counter()
        ctx.set_system(s)
        req.set_ms = 1e3 * (time.perf_counter() - req.start)
        try:
            info = ctx.build_grid()
        except RuntimeError as e:
            if "out of memory" in str(e) or "MiB free" in str(e):
                return None
            raise
        t1 = time.perf_counter()
        ctx.solve_async(s.rho_packed)
        E, F, inf = ctx.collect()
        req.solve_ms = 1e3 * (time.perf_counter() - t1)
        req.stage_ms = 1e3 * info["grid_seconds"]
        req.load_ms = 1e3 * info["upload_seconds"]
        req.load_phases = info["load_ms"]
        return int(info["arena_bytes"]), E, F, inf

    def _start(self, req: Request):
        r = req.replica
        s = self.replicas[r].take(req.step)
        if req.step < self.last_step[r]:
            self.replicas[r].prefetch(req.step + 1, self.pool)
        if self.ctx[r] is None:
            self.ctx[r] = g2g.Context(precision=self.precision, grid_mode=self.grid_mode)
            if self.density_variant is not None:
                self.ctx[r].set_schedule(density_variant=self.density_variant)
        if not self.resident[r]:
            self.arena[r] = self.arena[r] or self._estimate(r)
        self.resident[r] = True
        self.busy[r] = True
        return self.workers.submit(self._build, req, self.ctx[r], s)

    def _finish(self, req: Request, result, keep_fock: bool):
        r = req.replica
        arena, E, F, info = result
        self.arena[r] = arena
        req.energy, req.electrons = E, info["electrons"]
        req.fock = F if keep_fock else None
        req.done = time.perf_counter()
        self.last_use[r] = req.done
        self.busy[r] = False

    def run(self, arrivals: list[tuple[float, int]], keep_fock: bool = False,
            max_seconds: float = 600.0) -> dict:
        """Serves requests arriving at (time offset s, replica) in order; a
        replica's requests are its consecutive MD steps, continuing from the
        previous run on this server.  Returns the completed requests and the
        latency / throughput figures."""
        step = list(self.next_step)
        ev0, rt0 = self.evictions, self.retries
        reqs = []
        for i, (t, r) in enumerate(arrivals):
            step[r] += 1
            reqs.append(Request(i, r, step[r], t))
        self.last_step = list(step)
        self.next_step = list(step)
        pending: list[Request] = []
        inflight: dict = {}
        done: list[Request] = []
        t0 = time.perf_counter()
        nxt = 0
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            if now > max_seconds:
                raise RuntimeError("serve loop exceeded %.0f s" % max_seconds)
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                reqs[nxt].arrival += t0
                pending.append(reqs[nxt])
                nxt += 1
            progressed = False
            for fut in [f for f in inflight if f.done()]:  # collect finished builds
                req = inflight.pop(fut)
                res = fut.result()
                progressed = True
                if res is None:  # out of device memory: back in the queue, a larger estimate
                    r = req.replica
                    self.busy[r] = False
                    self.resident[r] = False
                    if not inflight and not any(self.resident):
                        raise RuntimeError("replica %d does not fit the device alone" % r)
                    self.arena[r] = int(1.5 * max(self.arena[r], self._estimate(r)))
                    self.retries += 1
                    pending.append(req)
                    pending.sort(key=lambda q: q.rid)
                    continue
                self._finish(req, res, keep_fock)
                done.append(req)
            for req in list(pending):   # admit: slot, replica idle, in order, memory
                if len(inflight) >= self.slots:
                    break
                r = req.replica
                if self.busy[r] or any(p.replica == r and p.rid < req.rid for p in pending):
                    continue
                if not self._make_room(r):
                    continue
                pending.remove(req)
                inflight[self._start(req)] = req
                progressed = True
            if not progressed:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        lat = np.array([1e3 * (q.done - q.arrival) for q in done])
        svc = np.array([1e3 * (q.done - q.start) for q in done])
        done.sort(key=lambda q: q.rid)
        return {"requests": done, "wall_s": wall, "builds_per_s": len(done) / wall,
                "latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)),
                               "mean": float(lat.mean())},
                "service_ms": {"p50": float(np.percentile(svc, 50)),
                               "p99": float(np.percentile(svc, 99))},
                "phase_ms_p50": {k: float(np.median([getattr(q, k + "_ms") for q in done]))
                                 for k in ("set", "stage", "load", "solve")},
                "load_phase_ms_p50": [float(np.median([q.load_phases[i] for q in done]))
                                      for i in range(4)],
                "evictions": self.evictions - ev0, "retries": self.retries - rt0}

    def close(self):
        self.workers.shutdown(wait=True)
        self.pool.shutdown(wait=True)
        for c in self.ctx:
            if c is not None:
                c.close()
        self.ctx = [None] * len(self.replicas)


def poisson_arrivals(n: int, rate_hz: float, nreplicas: int, seed: int) -> list[tuple[float, int]]:
    """n requests of a seeded Poisson process at rate_hz over nreplicas
    replicas chosen uniformly."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=n))
    r = rng.integers(0, nreplicas, size=n)
    return [(float(a), int(b)) for a, b in zip(t, r)]


def replica_set(n: int = 128, seed: int = 4, atoms=(10, 200)) -> list[Replica]:
    """n replicas whose QM regions are random molecules of ``atoms`` atoms,
    generated like the molecule batches (seeded per replica)."""
    mols = systems.molecule_batch(n, seed=seed, atoms=atoms)
    return [Replica(m, seed=1000 * seed + k) for k, m in enumerate(mols)]
