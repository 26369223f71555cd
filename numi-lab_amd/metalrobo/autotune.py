"""Launch-schedule autotuner for heterogeneous batches (worlds of different
trees and contact counts in one launch).

It decides the two knobs of the work-queue kernel:
  * tile assignment: which world a workgroup takes next. "natural" is index
    order; "heavy_first" orders worlds by their measured cost (shader cycles of
    one profiled control step), longest first, so that the slowest worlds start
    early and the tail of the launch shrinks;
  * wave assignment: how many persistent workgroups run (the GPU's resident
    capacity for this world, or a fraction of it).
plus the plain schedule (one workgroup per world, hardware dispatch order).
Each candidate is timed on graph-captured horizons of the world's own
commands, from the same saved state; the fastest is applied and reported.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


def snapshot(world):
    """State and persistent manifolds of every environment."""
    return (*world.state(), *world.manifolds())


def restore(world, snap) -> None:
    q, v, s, h, p, c = snap
    world.set_state(q, v, s)
    world.write_manifolds(h, p, c)


def measure_costs(world, commands) -> np.ndarray:
    """Shader cycles per environment over the control steps in `commands` (host array
    [steps][E][max_nv]) with the per-world schedule; the world's state is restored after."""
    snap = snapshot(world)
    world.set_schedule(None, 0)
    world.set_stage_profiling(True)
    world.advance(commands)
    cyc = world.stage_cycles().astype(np.float64).sum(1)
    world.set_stage_profiling(False)
    restore(world, snap)
    return cyc


@dataclass
class Choice:
    name: str
    order: np.ndarray | None
    grid: int
    ms_per_step: float = float("inf")


@dataclass
class TuneResult:
    best: Choice
    trials: list = field(default_factory=list)
    resident_workgroups: int = 0


def candidates(world, costs: np.ndarray) -> list:
    E = world.environment_count
    resident = world.resident_workgroups()
    natural = np.arange(E, dtype=np.uint32)
    heavy = np.argsort(-costs, kind="stable").astype(np.uint32)
    out = [Choice("grid", None, 0)]
    for frac in (1.0, 0.5):
        g = max(1, min(E, int(resident * frac)))
        out.append(Choice(f"queue_natural_{g}", natural, g))
        out.append(Choice(f"queue_heavy_first_{g}", heavy, g))
    return out


def autotune(world, commands_device, trial_horizons: int = 2, profile_commands=None) -> TuneResult:
    """Times every candidate schedule on `trial_horizons` replays of a graph-captured
    horizon over `commands_device` (torch tensor [steps][E][max_nv] at a fixed address) from
    the world's current state, applies the fastest and returns all timings. The world's
    state is restored afterwards."""
    import torch

    steps = int(commands_device.shape[0])
    if profile_commands is None:
        profile_commands = commands_device[:1].cpu().numpy()
    costs = measure_costs(world, profile_commands)
    snap = snapshot(world)
    choices = candidates(world, costs)
    trials = []
    for c in choices:
        world.set_schedule(c.order, c.grid)
        restore(world, snap)
        world.enqueue_graph(commands_device)  # capture (and first replay, untimed)
        world.wait()
        times = []
        for _ in range(trial_horizons):
            restore(world, snap)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            world.enqueue_graph(commands_device)
            world.wait()
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3 / steps)
        c.ms_per_step = float(min(times))
        trials.append({"name": c.name, "grid": c.grid, "ms_per_step": round(c.ms_per_step, 4)})
    best = min(choices, key=lambda c: c.ms_per_step)
    world.set_schedule(best.order, best.grid)
    restore(world, snap)
    return TuneResult(best, trials, world.resident_workgroups())
