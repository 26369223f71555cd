"""One-GPU step server: rollout requests (N worlds for H control steps) arrive
over time and are served from one resident pool of worlds kept in HBM.

Admission: a request takes N free slots of the pool when they exist and the
device memory of all admitted requests' worlds, at the measured bytes per
slot, stays within the HBM budget; otherwise it waits (first come, first
served). A request whose own worlds exceed the budget or the pool is refused. Its slots start from the
initial state through the control-step kernel's reset mask. Every tick
advances all active slots by one control step in a single launch of the
work-queue kernel over exactly those slots (inactive slots are not touched),
then retires requests whose horizon is complete and frees their slots.
Latency is measured per tick (every active request sees one control step per
tick); throughput is env-steps over the wall time of the loop.

The baseline is the same requests served one after another, each as its own
world with one plain launch per control step.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    worlds: int
    horizon: int
    arrival: float  # seconds after the server starts
    slots: np.ndarray | None = None
    done: int = 0
    admitted: float = -1.0
    finished: float = -1.0
    first_tick: int = -1  # the tick of its first control step (its command-bank row is first_tick % bank)


def poisson_requests(rate: float, count: int, seed: int = 0, worlds=(1, 1024), horizon=(10, 500)) -> list:
    """`count` requests with exponential inter-arrival times at `rate` per second, N
    log-uniform in `worlds`, horizon uniform in `horizon` (inclusive)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(worlds[0]), math.log(worlds[1])))))
        h = int(rng.integers(horizon[0], horizon[1] + 1))
        out.append(Request(i, max(worlds[0], min(worlds[1], n)), h, t))
    return out


@dataclass
class ServeReport:
    requests: int
    env_steps: int
    wall_s: float
    env_steps_per_s: float
    latency_ms_p50: float
    latency_ms_p99: float
    ticks: int
    mean_active_worlds: float
    max_active_worlds: int
    queue_wait_s_p50: float
    queue_wait_s_p99: float
    capacity_slots: int
    bytes_per_slot: float
    hbm_budget_bytes: float
    max_admitted_bytes: float = 0.0
    budget_blocked_ticks: int = 0  # ticks on which an arrived request waited on the budget with enough free slots
    tick_ms: list = field(default_factory=list, repr=False)
    results: dict = field(default_factory=dict, repr=False)


class StepServer:
    """Serves requests on a pool of `model.environment_count` resident worlds."""

    def __init__(self, model, config=None, hbm_budget_bytes: float = 200e9, command_bank: int = 32) -> None:
        import torch

        from .native import World

        self.model = model
        self.world = World(model, config)
        E = model.environment_count
        layout = self.world.layout()
        self.bytes_per_slot = float(layout.retained_device_bytes) / E
        self.set_hbm_budget(hbm_budget_bytes)
        q0, v0, s0 = model.initial_state()
        self.world.set_reset_state(q0, v0, s0)
        self.grid = self.world.resident_workgroups()
        # a bank of command sets on the device, cycled by tick
        self.commands = torch.from_numpy(model.commands(0, command_bank)).to("cuda")
        self.mask_host = np.zeros((1, E), np.uint32)
        self.mask = torch.zeros((1, E), dtype=torch.int32, device="cuda")

    def set_hbm_budget(self, hbm_budget_bytes: float) -> None:
        """Device-memory budget of the admitted worlds; capacity is the number of slots it covers."""
        self.hbm_budget = float(hbm_budget_bytes)
        self.capacity = int(min(self.model.environment_count, self.hbm_budget // max(self.bytes_per_slot, 1.0)))

    def serve(self, requests, idle_sleep: bool = True, return_states: bool = False) -> ServeReport:
        """Serves `requests` to completion. With `return_states`, the final (q, v, scene) rows of
        each request's worlds are read back when it retires (report.results[rid]); their slots
        are then free for later requests."""
        import torch

        too_big = [r.rid for r in requests if r.worlds > self.capacity]
        if too_big:
            raise ValueError(f"requests {too_big} need more worlds than the pool and the HBM budget hold "
                             f"({self.capacity} slots)")
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        free = deque(range(self.model.environment_count))
        active = []
        tick_ms, active_counts = [], []
        env_steps = 0
        results = {}
        t0 = time.perf_counter()
        tick = 0
        changed = True  # the active set differs from the one the schedule holds
        admitted_bytes, max_admitted_bytes, budget_blocked = 0.0, 0.0, 0
        while pending or active:
            now = time.perf_counter() - t0
            admitted = []
            while pending and pending[0].arrival <= now and len(free) >= pending[0].worlds:
                r = pending[0]
                need = r.worlds * self.bytes_per_slot
                if admitted_bytes + need > self.hbm_budget:
                    budget_blocked += 1
                    break
                pending.popleft()
                r.slots = np.array([free.popleft() for _ in range(r.worlds)], np.uint32)
                r.admitted = now
                r.first_tick = tick
                admitted_bytes += need
                max_admitted_bytes = max(max_admitted_bytes, admitted_bytes)
                admitted.append(r)
            if not active and not admitted:
                if pending and idle_sleep:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            ts = time.perf_counter()
            active.extend(admitted)
            n_active = sum(r.worlds for r in active)
            if admitted or changed:
                # the schedule and reset mask travel only when the active set changes
                order = np.concatenate([r.slots for r in active])
                self.world.set_schedule(order, min(self.grid, len(order)))
                changed = False
            mask = None
            if admitted:
                self.mask_host[:] = 0
                for r in admitted:
                    self.mask_host[0, r.slots] = 1
                self.mask.copy_(torch.from_numpy(self.mask_host.view(np.int32)))
                mask = self.mask
            k = tick % self.commands.shape[0]
            self.world.enqueue_device(self.commands[k:k + 1], mask)
            self.world.wait()
            tick_ms.append((time.perf_counter() - ts) * 1e3)
            active_counts.append(n_active)
            env_steps += n_active
            tick += 1
            still = []
            end = time.perf_counter() - t0
            states = None
            for r in active:
                r.done += 1
                if r.done >= r.horizon:
                    r.finished = end
                    if return_states:
                        states = states if states is not None else self.world.state()
                        results[r.rid] = tuple(a[r.slots.astype(np.int64)].copy() for a in states)
                    free.extend(int(s) for s in r.slots)
                    admitted_bytes -= r.worlds * self.bytes_per_slot
                    changed = True
                else:
                    still.append(r)
            active = still
        wall = time.perf_counter() - t0
        # every request-step of a tick sees that tick's latency
        lat = np.repeat(np.array(tick_ms), np.array(active_counts)) if tick_ms else np.zeros(1)
        waits = np.array([r.admitted - r.arrival for r in requests]) if requests else np.zeros(1)
        return ServeReport(
            requests=len(requests), env_steps=int(env_steps), wall_s=wall, env_steps_per_s=env_steps / wall,
            latency_ms_p50=float(np.percentile(lat, 50)), latency_ms_p99=float(np.percentile(lat, 99)),
            ticks=tick, mean_active_worlds=float(np.mean(active_counts)) if active_counts else 0.0,
            max_active_worlds=int(max(active_counts)) if active_counts else 0,
            queue_wait_s_p50=float(np.percentile(waits, 50)), queue_wait_s_p99=float(np.percentile(waits, 99)),
            capacity_slots=self.capacity, bytes_per_slot=self.bytes_per_slot, hbm_budget_bytes=self.hbm_budget,
            max_admitted_bytes=max_admitted_bytes, budget_blocked_ticks=budget_blocked,
            tick_ms=tick_ms, results=results)


def serve_serially(make_model, requests, config_of=None) -> dict:
    """The baseline: each request alone, as its own world of `worlds` environments, one plain
    launch per control step, one request after another. Returns env-steps, stepping time and
    per-step latencies (world construction excluded)."""
    import torch

    from .native import World

    env_steps, busy, lat = 0, 0.0, []
    for r in requests:
        model = make_model(r.worlds)
        cfg = config_of(model) if config_of is not None else model.default_config()
        w = World(model, cfg)
        cmds = torch.from_numpy(model.commands(0, min(r.horizon, 32))).to("cuda")
        torch.cuda.synchronize()
        for s in range(r.horizon):
            ts = time.perf_counter()
            w.enqueue_device(cmds[s % cmds.shape[0]:s % cmds.shape[0] + 1])
            w.wait()
            dt = time.perf_counter() - ts
            busy += dt
            lat.append(dt * 1e3)
        env_steps += r.worlds * r.horizon
        w.close()
    lat = np.array(lat) if lat else np.zeros(1)
    return {"requests": len(requests), "env_steps": int(env_steps), "stepping_s": busy,
            "env_steps_per_s": env_steps / busy if busy > 0 else 0.0,
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99))}
