"""Tile-shape autotuner.

The fused step's tile shape (cells per tile along x, y, z) sets the work of
one workgroup in the rotation pass, the departure lists and the per-substep
kernels, and with it the LDS footprint and the occupancy.  Results do not
depend on it (the tile re-sort, the per-cell sums in id order and the event
ranking are shape-independent), so the tuner may pick any candidate by
speed alone.  It times a few fused steps of the caller's system for each
candidate, from one common start-up state, and returns the fastest.
"""
from __future__ import annotations

import time

import numpy as np

from . import BW_FP64, System

MAX_CAP = 4096
AXIS_MAX = 64
TILE_REACH = 3


def tile_slots(ncell: int, density: int) -> int:
    """Slots per tile as the library sizes them (mean + 15 % + 64, to 64)."""
    need = (ncell * density * 23 + 19) // 20 + 64
    return (need + 63) // 64 * 64


def candidates(L: int, density: int, fp64: bool = False) -> list[tuple[int, int, int]]:
    """Tile shapes the library accepts for a box of edge L: each edge divides
    L and is at least 4, at most 512 cells, the slots and the SRD pass's LDS
    (per-cell FP64 sums, per-slot staging) fit, the region fits the masks.
    The SRD workgroup follows the slots (3 or 4 per thread, 256, 512 or
    1024 threads: the fewest slots that hold the tile), so smaller tiles let
    several workgroups share a CU."""
    divs = [b for b in range(4, min(L, 16) + 1) if L % b == 0] or [L]
    rb = 8 if fp64 else 4
    out = []
    for bx in divs:
        for by in divs:
            for bz in divs:
                if by > bx or bz > by:
                    continue
                n = bx * by * bz
                cap = tile_slots(n, density)
                if n < 64 or n > 512 or cap > MAX_CAP:
                    continue
                if any(min(b + 2 * TILE_REACH, L) > AXIS_MAX for b in (bx, by, bz)):
                    continue
                lds = max(n * 56, cap * 4) + 3 * cap * rb + (2 * n + 1) * 4 + cap * 2
                if lds > 160 * 1024 - 16 * 1024:
                    continue
                out.append((bx, by, bz))
    return out


def autotune(params: dict, steps: int = 4, warmup: int = 2, shapes=None, v0=None, log=None) -> dict:
    """Time `steps` fused steps for each candidate tile shape of one system
    (params as for System) from one common start-up state; returns
    {"best": shape, "ms_per_step": {shape: ms}}."""
    prm = {k: v for k, v in params.items() if k != "tile"}
    base = System(**prm)
    base.startup()
    pos, vel, _ = base.download_fluid()
    col = base.download_colloid()
    base.close()
    fp64 = prm.get("precision", 0) == BW_FP64
    shapes = shapes or candidates(prm["L"], prm.get("density", 10), fp64)
    times = {}
    for shape in shapes:
        t0 = time.time()
        s = System(tile=shape, **prm)
        s.initialize()
        s.upload_fluid(pos, vel)
        s.upload_colloid(**col)
        s.neighbour_list_md()
        s.compute_force_md()
        if v0 is not None:
            s.set_v0(v0)
        s.step(warmup)
        s.status()
        ms = s.time_steps(steps) / steps
        s.status()
        s.close()
        times[shape] = ms
        if log:
            log(f"tile {shape}: {ms:.3f} ms/step ({time.time() - t0:.1f} s)")
    best = min(times, key=times.get)
    return {"best": best, "ms_per_step": times}


def common_candidates(Ls, density: int = 10, fp64: bool = False) -> list[tuple[int, int, int]]:
    """Tile shapes every box edge of an ensemble accepts (one slot capacity
    and one SRD workgroup size serve all its replicas)."""
    sets = [set(candidates(L, density, fp64)) for L in sorted(set(Ls))]
    return sorted(set.intersection(*sets), key=lambda s: (s[0] * s[1] * s[2], s))


def _time_ensemble(replicas: list, tiles: dict, steps: int, warmup: int) -> float:
    """ms per fused step of the ensemble with tile shape tiles[L] for each box edge L."""
    s = System(replicas=[dict(r, tile=tiles[r["L"]]) for r in replicas])
    s.startup()
    s.step(warmup)
    s.status()
    ms = s.time_steps(steps) / steps
    s.status()
    s.close()
    return ms


def autotune_ensemble(replicas: list, steps: int = 5, warmup: int = 2, shapes=None, per_box: bool = False,
                      max_cells: int = 256, even: bool = False, passes: int = 2, log=None) -> dict:
    """Time `steps` fused steps of the whole ensemble (replica params as for
    System(replicas=...)) for each tile shape its box edges all accept;
    returns {"best": shape, "ms_per_step": {shape: ms}}.

    per_box: then refine the shape per box edge, one edge at a time from the
    best common shape (every shape of at most max_cells cells that edge
    accepts, the others held), keeping a change that is faster by more than
    1 % (the timing's spread), in passes over the edges until one changes
    nothing (at most `passes`: a change can pay only once another has been
    made, when the largest tile sets every replica's slot capacity); adds
    {"per_box": {L: shape}, "per_box_ms": ms, "per_box_tried": {"L:shape": ms}}.  The replicas then share one slot
    capacity and SRD workgroup size (the largest tile's), which the timing
    includes.  even: only shapes of even edges (the surrogate closure's 2x2x2
    coarse cells need them)."""
    fp64 = replicas[0].get("precision", 0) == BW_FP64
    density = replicas[0].get("density", 10)
    shapes = shapes or common_candidates([r["L"] for r in replicas], density, fp64)
    if even:
        shapes = [s for s in shapes if not any(e % 2 for e in s)]
        if not shapes:
            raise ValueError("no tile shape with even edges divides every box edge of the ensemble")
    times = {}
    for shape in shapes:
        t0 = time.time()
        times[shape] = _time_ensemble(replicas, {r["L"]: shape for r in replicas}, steps, warmup)
        if log:
            log(f"ensemble tile {shape}: {times[shape]:.3f} ms/step ({time.time() - t0:.1f} s)")
    best = min(times, key=times.get)
    out = {"best": best, "ms_per_step": times}
    if per_box:
        Ls = sorted(set(r["L"] for r in replicas))
        assign = {L: best for L in Ls}
        best_ms = times[best]
        tried = {}
        for _ in range(passes):
            changed = False
            for L in Ls:
                for shape in candidates(L, density, fp64):
                    if shape == assign[L] or shape[0] * shape[1] * shape[2] > max_cells:
                        continue
                    if even and any(e % 2 for e in shape):
                        continue
                    trial = dict(assign)
                    trial[L] = shape
                    ms = _time_ensemble(replicas, trial, steps, warmup)
                    tried[f"{L}:{shape}"] = ms
                    if log:
                        log(f"ensemble tile {shape} for L={L}: {ms:.3f} ms/step")
                    if ms < 0.99 * best_ms:
                        assign, best_ms, changed = trial, ms, True
            if not changed:
                break
        out.update(per_box=assign, per_box_ms=best_ms, per_box_tried=tried)
    return out
