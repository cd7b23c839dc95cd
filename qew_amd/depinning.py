"""Depinning threshold search over many disorder realizations at once.

The reference's ramp (src/qew_runs.h:224-241) brackets f_c between the last
blocked and the first unblocked force of f = k * f_step; its only bisection
(tests/test_rosso_krauth.cpp:76-88) halves a bracket one rk_block probe at a
time, a timeout counting as not blocked. fc_search does the same decisions
with k_way forces per line per round: each round probes
lo + j (hi - lo) / (k_way + 1), j = 1..k_way, every probe started from the
line's highest blocked configuration (no-passing), all probes of all lines in
one batched rk_block. The new bracket runs from the highest force below
which every probe blocked to the next probe, so a round shrinks the bracket
k_way + 1 times instead of 2.
"""
import time

import numpy as np

from . import Batch, Model, Params


def fc_search(noises, n, y0, rp, lengths_L=None, line_tension=1.0, model=Model.Linear,
              f_start=0.125, k_way=4, rel_tol=1e-10, budget_s=float("inf"), log=None,
              record=None, state=None, checkpoint=None):
    """f_c bracket of every potential in `noises` for a line of n sites (box
    length lengths_L, default n) started flat at y0: relaxation at f = 0,
    ramp f = f_start * 2^k per line until it is not blocked, then k_way-way
    rounds until every bracket is within rel_tol (relative) or the budget
    is spent. Returns (lo, hi, h_lo, stats): brackets, the blocked
    configurations at lo (torch, nlines x n) and counters. `record(k, res, h,
    seconds)` sees the relaxation (k = 0) and the first ramp probe (k = 1).
    `state` = (lo, hi, h_lo) of an earlier call continues its rounds (no
    relaxation or ramp); `checkpoint(lo, hi, h_lo)` is called after the ramp
    and after every round with the state such a call continues from."""
    import torch
    nlines = len(noises)
    L = float(n if lengths_L is None else lengths_L)
    dev = torch.device("cuda:0")
    h_lo = torch.full((nlines, n), float(y0), dtype=torch.float64, device=dev)
    lo = np.zeros(nlines)
    hi = np.full(nlines, np.inf)
    stats = {"rounds": 0, "probes": 0, "batch_sweeps": 0, "line_sweeps": 0, "evals_done": 0,
             "evals_reference": 0, "round_log": []}
    t0 = time.perf_counter()

    def run(rows, forces):
        b = Batch([n] * len(rows), [noises[r] for r in rows],
                  [Params(L, line_tension, 0.0, model)] * len(rows))
        b.set_forces(forces)
        h = h_lo[torch.as_tensor(rows, device=dev)].reshape(-1).clone()
        t = time.perf_counter()
        res = b.rk_block(h, rp)
        sec = time.perf_counter() - t
        stats["probes"] += len(rows)
        stats["batch_sweeps"] += max(r.sweeps for r in res)
        stats["line_sweeps"] += sum(r.sweeps for r in res)
        stats["evals_done"] += sum(r.n_velocity_evals_done for r in res)
        stats["evals_reference"] += sum(r.n_velocity_evals for r in res)
        return res, h.view(len(rows), n), sec

    k = 0
    f_ramp = f_start
    forces = np.zeros(nlines)
    if state is not None:
        lo[:] = state[0]
        hi[:] = state[1]
        h_lo.copy_(torch.as_tensor(np.asarray(state[2]), device=dev).view(nlines, n))
    while state is None:
        rows = np.nonzero(np.isinf(hi))[0] if k else np.arange(nlines)
        res, h, sec = run(rows, forces[rows])
        if record and k < 2:
            record(k, res, h, sec)
        blocked = np.array([r.blocked for r in res])
        if k == 0 and not blocked.all():
            raise RuntimeError("a line does not block at f = 0")
        sel = np.nonzero(blocked)[0]
        h_lo[torch.as_tensor(rows[sel], device=dev)] = h[torch.as_tensor(sel, device=dev)]
        lo[rows[blocked]] = forces[rows[blocked]]
        hi[rows[~blocked]] = forces[rows[~blocked]]
        k += 1
        if not np.isinf(hi).any():
            break
        forces = np.where(np.isinf(hi), np.maximum(2 * lo, f_ramp), forces)
        f_ramp *= 2
    stats["ramp_probes"] = k
    if checkpoint and state is None:
        checkpoint(lo, hi, h_lo)
    while True:
        todo = np.nonzero(hi - lo > rel_tol * np.maximum(np.abs(lo), np.abs(hi)))[0]
        if todo.size == 0 or time.perf_counter() - t0 > budget_s:
            break
        rows = np.repeat(todo, k_way)
        j = np.tile(np.arange(1, k_way + 1), todo.size)
        f = lo[rows] + j * (hi[rows] - lo[rows]) / (k_way + 1)
        res, h, sec = run(rows, f)
        blocked = np.array([r.blocked for r in res]).reshape(todo.size, k_way)
        sweeps = np.array([r.sweeps for r in res]).reshape(todo.size, k_way)
        f = f.reshape(todo.size, k_way)
        for t, line in enumerate(todo):
            jb = 0  # probes 1..jb all blocked
            while jb < k_way and blocked[t, jb]:
                jb += 1
            if jb > 0:
                lo[line] = f[t, jb - 1]
                h_lo[line] = h[t * k_way + jb - 1]
            if jb < k_way:
                hi[line] = f[t, jb]
        stats["rounds"] += 1
        width = (hi - lo) / np.maximum(np.abs(hi), 1e-300)
        stats["round_log"].append({"seconds": round(sec, 3), "lines": int(todo.size),
                                   "max_sweeps": int(sweeps.max()),
                                   "rel_width_max": float(width.max())})
        if log:
            log(f"  round {stats['rounds']}: {todo.size} lines x {k_way}, {sec:.2f} s, "
                f"max sweeps {sweeps.max()}, max rel width {width.max():.3g}")
        if checkpoint:
            checkpoint(lo, hi, h_lo)
    stats["seconds"] = time.perf_counter() - t0
    return lo, hi, h_lo, stats
