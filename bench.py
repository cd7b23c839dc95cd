"""Benchmark of the qEW gradient/objective hot path on one MI355X.

Workload (the gradient/objective throughput configuration): a line of
L = 65536 sites on a 65536 x 8192 FilteredNoise table (xi = 8 grid units,
unit amplitude, seed 1, grid units dx = dy = 1), arclength elasticity,
line tension 1, no driving force. The line is a seeded random walk with
unit-variance Gaussian steps started at Nh/2. One step = 10^4 fused
gradient + energy evaluations issued as one batched launch; evaluation r sees
the walk shifted by 520 r cells (mod Nh), more than the walk's 500-cell band,
so the evaluations in flight read disjoint table rows, and the 1024 distinct
shifts cycle through the whole table (no L2/MALL reuse across evaluations).
Line states and gradients of all 10^4 evaluations are distinct buffers in HBM.

`value` is FP32-compensated evaluations per second (FP32 line offsets from a
per-line anchor, FP32 table, two-sum energy); the FP64 path (bit-exact to the
oracle) is reported beside it. The roofline uses the compulsory bytes per
evaluation, s_c * sum_r U_r + L * (s_h + s_g), with U_r the number of distinct
coefficients x-row r supplies (computed exactly from the walk).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
L = 65536
NY = 8192
XI = 8.0
EVALS = 10000
# evaluation r sees the walk shifted by (r * SHIFT) mod NY cells; the walk's
# band (its y-range, about 500 cells) is narrower than SHIFT, so evaluations in
# flight together read disjoint table rows
SHIFT = 520
# Rosso-Krauth estimate of the depinning force of the Langevin workload's
# disorder (65536 x 4096 table, xi = 4, seed 1, linear, line tension 1, FP32
# table, the depinning tool's knobs): blocked at 0.5078, not blocked at
# 0.5156 (a 2*10^5-sweep timeout); ramp and bisection as in bench_rk.
FC_EST_LANGEVIN = 0.512

ROOT = os.path.dirname(os.path.abspath(__file__))


_ALL_CPUS = None


def omp_setup():
    """Thread placement of the OpenMP CPU baselines, set before torch or the
    oracle loads the OpenMP runtime: one thread per core this process may run
    on (an OMP_NUM_THREADS already set is kept), spread over the cores. An
    OpenMP runtime that reads these binds the thread that loads it to one
    core; unpin() gives that thread its whole CPU set back, so that the host
    threads the library starts (table generation and conversion) are not
    confined to one core."""
    global _ALL_CPUS
    _ALL_CPUS = os.sched_getaffinity(0)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    nproc = len(_ALL_CPUS)
    os.environ.setdefault("OMP_NUM_THREADS", str(nproc))
    return {"nproc": nproc, "OMP_NUM_THREADS": os.environ["OMP_NUM_THREADS"],
            "OMP_PROC_BIND": os.environ["OMP_PROC_BIND"], "OMP_PLACES": os.environ["OMP_PLACES"]}


def unpin():
    if _ALL_CPUS:
        os.sched_setaffinity(0, _ALL_CPUS)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    O.load(omp=True)
    unpin()
    return O


def walk(seed=2):
    rng = np.random.default_rng(seed)
    return NY / 2 + np.cumsum(rng.standard_normal(L))


def distinct_coefficients(h, n, nx, ny):
    """sum_r U_r: distinct (x-row, y-cell) pairs the edges of line h touch."""
    i = np.arange(n)
    ip = (i + 1) % n
    gx = (i + ip) * 0.5 * (nx / n)
    ix = np.floor(gx).astype(np.int64) % nx
    hc = (h[ip] + h) * 0.5
    iy = np.floor(np.mod(hc, ny)).astype(np.int64)
    keys = []
    for a in (-1, 0, 1, 2):
        for b in (-1, 0, 1, 2):
            keys.append(((ix + a) % nx) * ny + (iy + b) % ny)
    return int(np.unique(np.concatenate(keys)).size)


def distinct_mid_values(h, n, nx, ny):
    """Distinct (x-column, y-cell) values a midpoint-contracted table supplies
    to the edges of line h (line spacing = table spacing): edge i reads the 4
    y levels of column (i + ip) // 2; only the wrap edge shares a column."""
    i = np.arange(n)
    ip = (i + 1) % n
    ix = ((i + ip) // 2) % nx
    hc = (h[ip] + h) * 0.5
    iy = np.floor(np.mod(hc, ny)).astype(np.int64)
    keys = [ix * ny + (iy + b) % ny for b in (-1, 0, 1, 2)]
    return int(np.unique(np.concatenate(keys)).size)


def edge_cells(h, n, nx, ny):
    """(x cell, y cell) of every edge midpoint of line h (line spacing = table
    spacing): edge i reads x-rows ix-1..ix+2 and y-rows iy-1..iy+2."""
    i = np.arange(n)
    ip = (i + 1) % n
    gx = (i + ip) * 0.5 * (nx / n)
    ix = np.floor(gx).astype(np.int64) % nx
    hc = (h[ip] + h) * 0.5
    iy = np.floor(np.mod(hc, ny)).astype(np.int64)
    return ix, iy


def count_lines(h, n, nx, ny, value_bytes=4, tile=None, mid=False):
    """Distinct 128-byte lines the table gathers of one evaluation touch.
    tile=None: x-contiguous rows (the device layout); tile=(tx, ty): blocks of
    tx x-columns by ty y-rows per 128-byte line, x fastest inside a block.
    mid: one value per edge column (midpoint-contracted table), else the
    4 x 4 window."""
    ix, iy = edge_cells(h, n, nx, ny)
    per_line = 128 // value_bytes
    xs = [0] if mid else [-1, 0, 1, 2]
    keys = []
    for a in xs:
        x = (ix + a) % nx
        for b in (-1, 0, 1, 2):
            y = (iy + b) % ny
            if tile is None:
                keys.append(y * ((nx + per_line - 1) // per_line) + x // per_line)
            else:
                tx, ty = tile
                keys.append((y // ty) * ((nx + tx - 1) // tx) + x // tx)
    return int(np.unique(np.concatenate(keys)).size)


def layout_report(h, n, nx, ny):
    """Lines per evaluation of the benchmark walk for the candidate layouts;
    bytes over the compulsory table bytes (s_c sum_r U_r, or 4 values per
    edge for the midpoint table)."""
    out = {}
    comp4 = 4 * distinct_coefficients(h, n, nx, ny)
    compm = 4 * distinct_mid_values(h, n, nx, ny)
    for name, tile in (("x-contiguous", None), ("16x2", (16, 2)), ("8x4", (8, 4)),
                       ("4x8", (4, 8))):
        for kind, mid, comp in (("fp32_4x4", False, comp4), ("fp32_mid", True, compm)):
            lines = count_lines(h, n, nx, ny, 4, tile, mid)
            out[f"{kind}/{name}"] = {"lines": lines, "bytes_over_compulsory": 128 * lines / comp}
    lines = count_lines(h, n, nx, ny, 8, None, False)
    out["fp64_4x4/x-contiguous"] = {"lines": lines,
                                    "bytes_over_compulsory": 128 * lines / (2 * comp4)}
    return out


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "eval_traffic.json")
    if os.path.exists(p):
        try:
            return json.load(open(p))
        except Exception:
            return None
    return None


def cpu_baseline(coef_rowmajor, h0, budget_s=15.0):
    """The oracle's OpenMP build (restatement of the reference's two-pass
    gradient + separate objective reduce) on this host, on a bounded sample of
    the same evaluations."""
    O = oracle()
    lib = O.load(omp=True)
    nz = O.OracleNoise.table(coef_rowmajor, float(L), float(NY))
    p = O.params(float(L), 1.0, 0.0, O.ARCLENGTH)
    t0 = time.perf_counter()
    k = 0
    while True:
        h = h0 + float((SHIFT * k) % NY)
        O.gradient(h, p, nz, omp=True)
        O.objective(h, p, nz, omp=True)
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": k / dt, "unit": "grad evals/s", "cores": int(lib.orc_num_threads()),
            "kind": "port",
            "sample": f"{k} gradient+objective evaluations (FP64, L={L}, {L}x{NY} table, "
                      f"OpenMP restatement of the reference two-pass gradient) in {dt:.1f} s"}


def bench_lbfgs(log, nlines=512, plain_sample=8, with_cpu=True, omp_info=None):
    """Batched static minimisation: nlines independent realizations with line
    lengths drawn from {1024, ..., 16384}, each on its own L x 2048 table with
    xi from {2, 4, 8}, linear or arclength elasticity (seeded draw), flat start
    at Ly/2, line tension A (lambda_p/xi)^{3/2} with lambda_p/xi = 4, L-BFGS
    m = 10 with the FFT preconditioner and the drivers' per-line shift
    0.5 dx A / xi^2. One persistent launch per precision; the plain FP64
    host-driven single-line path is timed on a sample of the same lines."""
    import torch
    import qew_amd as q
    rng = np.random.default_rng(2024)
    lengths = rng.choice([1024, 2048, 4096, 8192, 16384], nlines).tolist()
    xis = rng.choice([2.0, 4.0, 8.0], nlines).tolist()
    models = rng.integers(0, 2, nlines).tolist()
    ny = 2048
    t0 = time.perf_counter()
    specs = [(L_, ny, float(L_), float(ny), 1.0, xi, xi, 1 + k)
             for k, (L_, xi) in enumerate(zip(lengths, xis))]
    n64 = q.FilteredNoise.generate_batch(specs, dtype=q.F64)
    log(f"lbfgs: {nlines} FP64 tables generated in {time.perf_counter() - t0:.1f} s")
    n32 = [nz.convert(q.F32) for nz in n64]
    # every line has n = L sites on an L-wide table of spacing 1: its edge
    # midpoints lie at x-fraction 1/2, so midpoint-contracted tables apply
    nmid = [nz.convert(q.F32_MID) for nz in n64]
    log(f"lbfgs: FP32 and midpoint tables converted at {time.perf_counter() - t0:.1f} s")
    n64mid = [nz.convert(q.F64_MID) for nz in n64]
    torch.cuda.synchronize()
    log(f"lbfgs: {nlines} tables generated in {time.perf_counter() - t0:.1f} s "
        f"(sum L = {sum(lengths)})")
    lt = 4.0 ** 1.5
    params = [q.Params(float(L_), lt, 0.0, q.Model(m)) for L_, m in zip(lengths, models)]
    shifts = [0.5 / xi ** 2 for xi in xis]
    h0 = torch.full((sum(lengths),), ny / 2.0, dtype=torch.float64, device="cuda:0")
    out = {}
    finals, gmax, minima, results = {}, {}, {}, {}
    offs = np.concatenate([[0], np.cumsum(lengths)])
    m_hist = 10
    # ftol 1e-5: above the Armijo roundoff floor of the longest lines (L=16384,
    # |E| ~ 1e4), where the reference algorithm converges in ~100-300 iterations
    # FP32: the line offsets' resolution floors the FP32 gradient near 1e-4, so
    # FP32 stops there and is compared by its energy against the FP64 minimum
    for name, noises, dtype, ftol in (("f64", n64, q.F64, 1e-5), ("f32", n32, q.F32, 1e-4),
                                      ("f32_mid", nmid, q.F32, 1e-4),
                                      ("f64_mid", n64mid, q.F64, 1e-5),
                                      # FP64 stopped where FP32 stops, for a
                                      # like-for-like minima/s comparison
                                      ("f64_ftol1e-4", n64, q.F64, 1e-4)):
        b = q.Batch(lengths, noises, params)
        b.set_precond_shifts(shifts)
        lp = q.LbfgsParams(m=10, ftol=ftol, max_iter=10000, precondition=True)
        # untimed warm-up: the batch's first call picks its team policy by
        # timing the candidates on a copy of the line state
        h = h0.clone()
        torch.cuda.synchronize()
        t_tune = time.perf_counter()
        b.lbfgs(h, lp, dtype=dtype)
        t_tune = time.perf_counter() - t_tune
        wgs, team_arr = b.lbfgs_plan()
        team_count = {f"{L_}x{k_}": int(np.sum((np.array(lengths) == L_) & (team_arr == k_)))
                      for L_ in sorted(set(lengths)) for k_ in (1, 2, 4)
                      if np.any((np.array(lengths) == L_) & (team_arr == k_))}
        h = h0.clone()
        torch.cuda.synchronize()
        log(f"lbfgs {name}: launching (lines x workgroups {team_count}, picked in {t_tune:.2f} s)")
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t1 = time.perf_counter()
        ev0.record()
        res = b.lbfgs(h, lp, dtype=dtype)
        ev1.record()
        dt = time.perf_counter() - t1
        torch.cuda.synchronize()
        span = ev0.elapsed_time(ev1) * 1e-3
        # energies and gradients of the minima, evaluated in FP64 on the FP64 tables
        e, g = q.Batch(lengths, n64, params).eval(h)
        finals[name] = e.cpu().numpy()
        minima[name] = h.cpu().numpy()
        ga = g.abs().cpu().numpy()
        gmax[name] = np.array([ga[o0:o1].max() for o0, o1 in zip(offs[:-1], offs[1:])])
        its = [r.iterations for r in res]
        results[name] = (res, span)
        # single lines below run on as many workgroups as in the batch
        # the critical path: the line with the most site-iterations, solved
        # alone (on its team size in the batch), against the whole launch
        crit = int(np.argmax(np.array(its) * np.array(lengths)))
        os.environ["QEW_LBFGS_TEAM"] = str(int(team_arr[crit]))
        b1 = q.Batch([lengths[crit]], [noises[crit]], [params[crit]])
        b1.set_precond_shifts([shifts[crit]])
        h1 = torch.full((lengths[crit],), ny / 2.0, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        (r1,) = b1.lbfgs(h1, lp, dtype=dtype)
        alone = time.perf_counter() - t2
        del b1, h1
        # workgroup utilisation, estimated: every 8th line solved alone (on
        # as many workgroups as it has in the batch),
        # its workgroup-seconds scaled to the batch, over the launch's
        # workgroups x span (solo lines meet no contention, so this
        # underestimates the busy share somewhat)
        busy_sample = 0.0
        sample = list(range(0, nlines, 8))
        for k in sample:
            os.environ["QEW_LBFGS_TEAM"] = str(int(team_arr[k]))
            bk = q.Batch([lengths[k]], [noises[k]], [params[k]])
            bk.set_precond_shifts([shifts[k]])
            hk = torch.full((lengths[k],), ny / 2.0, dtype=torch.float64, device="cuda:0")
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            bk.lbfgs(hk, lp, dtype=dtype)
            busy_sample += (time.perf_counter() - t3) * bk.lbfgs_workgroups()
            del bk, hk
        busy = busy_sample * nlines / len(sample)
        del os.environ["QEW_LBFGS_TEAM"]
        out[name] = {"minima_per_s": nlines / dt, "seconds": dt, "ftol": ftol,
                     "launch_s": span,
                     "teams": {"lines_by_length_x_workgroups": team_count,
                               "picked_in_s": t_tune},
                     "critical_line": {"line": crit, "L": int(lengths[crit]),
                                       "workgroups": int(team_arr[crit]),
                                       "iterations": int(its[crit]), "alone_s": alone,
                                       "alone_over_launch": alone / span},
                     "converged": int(sum(r.converged for r in res)),
                     "curvature_skips": int(sum(r.n_curvature_skips for r in res)),
                     "ls_restarts": int(sum(r.n_ls_restarts for r in res)),
                     "slowest_line": int(np.argmax(its)),
                     "iterations_mean": float(np.mean(its)), "iterations_max": int(max(its)),
                     "objective_evals_per_iteration": float(
                         sum(r.n_func_evals for r in res) / max(1, sum(its))),
                     "critical_line_objective_evals_per_iteration": float(
                         res[crit].n_func_evals / max(1, its[crit])),
                     "workgroup_utilisation": {
                         "value": busy / (wgs * span), "workgroups": wgs,
                         "busy_workgroup_s_estimate": busy,
                         "sample": f"every 8th line solved alone ({len(sample)} lines)"}}
        log(f"lbfgs {name}: {json.dumps(out[name])}")
        del b, h
    # algorithmic bytes of each minimisation: per iteration the vector algebra
    # of the two-loop recursion and the update, (2m + 8) n s (history read
    # once); per function evaluation (line-search trials, n_func_evals) and per
    # gradient (iterations + 1) the distinct coefficients the line's edges
    # touch, s_c sum_r U_r (taken at the FP64 minimum), plus 2 n s of line
    # vectors; divided by the launch span
    u_line = np.array([distinct_coefficients(minima["f64"][o0:o1], o1 - o0, o1 - o0, ny)
                       for o0, o1 in zip(offs[:-1], offs[1:])], dtype=np.float64)
    u_mid = np.array([distinct_mid_values(minima["f64"][o0:o1], o1 - o0, o1 - o0, ny)
                      for o0, o1 in zip(offs[:-1], offs[1:])], dtype=np.float64)
    n_arr = np.array(lengths, dtype=np.float64)
    for name, (res, span) in results.items():
        s_state = 8 if name.startswith("f64") else 4
        s_c = 8 if name.startswith("f64") else 4
        u = u_mid if "mid" in name else u_line
        its = np.array([r.iterations for r in res], dtype=np.float64)
        nfe = np.array([r.n_func_evals for r in res], dtype=np.float64)
        per_eval = s_c * u + 2 * n_arr * s_state
        bytes_total = float((its * (2 * m_hist + 8) * n_arr * s_state
                             + (nfe + its + 1) * per_eval).sum())
        ach = bytes_total / span / 1e9
        out[name]["roofline"] = {"bound": "latency (the critical line's iteration chain, "
                                 "one workgroup or a team of them per line)", "achieved": ach,
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": ach / HBM_PEAK_GBS, "bytes": bytes_total}
    def energy_error(name):
        # relative energy difference per line against FP64; the lines beyond
        # 1e-6 are listed with the signed difference (negative: the other
        # precision's minimum is lower)
        d = finals[name] - finals["f64"]
        rel_e = np.abs(d) / np.abs(finals["f64"])
        return {"max": float(rel_e.max()), "median": float(np.median(rel_e)),
                "lines_within_1e-6": int((rel_e <= 1e-6).sum()),
                "beyond_1e-6": [{"line": int(k), "L": int(lengths[k]),
                                 "signed_rel": float(d[k] / abs(finals["f64"][k]))}
                                for k in np.flatnonzero(rel_e > 1e-6)]}
    out["f32_vs_f64_energy_rel_err"] = energy_error("f32")
    out["f64_mid_vs_f64_energy_rel_err"] = energy_error("f64_mid")
    out["f32_mid_vs_f64_energy_rel_err"] = energy_error("f32_mid")
    # max |grad| per line at each precision's minima, from the FP64 gradient
    out["max_grad_fp64_at_minima"] = {k: {"median": float(np.median(v)), "max": float(v.max())}
                                      for k, v in gmax.items()}
    # roughness of the minima: squared width per line, and the roughness
    # exponent from the height-difference ACF averaged over the lines of one
    # length (fit over 16 <= ell <= L/16 grid units), FP32 against FP64
    from qew_amd import analysis
    w2 = {k: np.array([analysis.width2(v[o0:o1]) for o0, o1 in zip(offs[:-1], offs[1:])])
          for k, v in minima.items()}
    w2_rel = np.abs(w2["f32"] - w2["f64"]) / w2["f64"]
    zeta = {}
    for L_ in sorted(set(lengths)):
        rows = [k for k in range(nlines) if lengths[k] == L_]
        z = {name: analysis.roughness_exponent(
                 np.stack([v[offs[k]:offs[k + 1]] for k in rows]), float(L_), 16.0, L_ / 16.0)
             for name, v in minima.items()}
        zeta[str(L_)] = {"f64": z["f64"], "f32": z["f32"], "abs_err": abs(z["f32"] - z["f64"]),
                         "lines": len(rows)}
    out["f32_vs_f64_roughness"] = {
        "W2_rel_err": {"median": float(np.median(w2_rel)), "max": float(w2_rel.max())},
        "zeta_by_length": zeta,
        "zeta_max_abs_err": max(v["abs_err"] for v in zeta.values())}
    # plain FP64 host-driven path, one line at a time, on a sample of lines
    idx = list(range(0, nlines, max(1, nlines // plain_sample)))[:plain_sample]
    t2 = time.perf_counter()
    for k in idx:
        h = torch.full((lengths[k],), ny / 2.0, dtype=torch.float64, device="cuda:0")
        q.lbfgs_minimize(h, params[k], n64[k],
                         q.LbfgsParams(m=10, ftol=1e-5, max_iter=10000, precondition=True,
                                       precond_shift=shifts[k]))
    tp = time.perf_counter() - t2
    plain = len(idx) / tp
    out["plain_f64_single_line_minima_per_s"] = plain
    out["plain_sample"] = (f"lines {idx} (evenly spaced in the batch; lengths "
                           f"{[lengths[k] for k in idx]}), {tp:.1f} s")
    out["speedup_f32_batched_vs_plain_f64"] = out["f32"]["minima_per_s"] / plain
    out["speedup_f64_batched_vs_plain_f64"] = out["f64"]["minima_per_s"] / plain
    out["speedup_f32_mid_batched_vs_plain_f64"] = out["f32_mid"]["minima_per_s"] / plain
    out["speedup_f64_mid_batched_vs_plain_f64"] = out["f64_mid"]["minima_per_s"] / plain
    # FP32 against FP64 at the same stopping rule (max|grad| < 1e-4)
    out["f32_vs_f64_at_same_ftol"] = {
        "ftol": 1e-4, "f64_minima_per_s": out["f64_ftol1e-4"]["minima_per_s"],
        "f32_minima_per_s": out["f32"]["minima_per_s"],
        "f32_mid_minima_per_s": out["f32_mid"]["minima_per_s"],
        "f64_iterations_mean": out["f64_ftol1e-4"]["iterations_mean"],
        "f32_iterations_mean": out["f32"]["iterations_mean"]}
    out["workload"] = (f"{nlines} lines, L from {{1024..16384}} (sum {sum(lengths)}), "
                       f"tables L x {ny}, xi in {{2,4,8}}, m=10, FFT preconditioner")
    if with_cpu:
        out["cpu_baseline"], out["f64_vs_oracle"] = lbfgs_oracle_sample(
            log, n64, lengths, xis, models, lt, shifts, ny, minima["f64"], offs,
            results["f64"][0], omp_info)
    log(f"lbfgs: {json.dumps(out)}")
    return out


def lbfgs_oracle_sample(log, n64, lengths, xis, models, lt, shifts, ny, h64, offs, res64,
                        omp_info, stride=2):
    """The oracle's L-BFGS (OpenMP over lines, each line sequential) on every
    stride-th line of the batch: the CPU baseline in minima/s, and the GPU FP64
    batch's minima against the oracle's (energies from the sequential oracle
    objective on the same table)."""
    O = oracle()
    idx = list(range(0, len(lengths), stride))
    t0 = time.perf_counter()
    tables = [O.OracleNoise.table(n64[k].coefficients(), float(lengths[k]), float(ny))
              for k in idx]
    log(f"lbfgs cpu: {len(idx)} tables copied in {time.perf_counter() - t0:.1f} s")
    po = [O.params(float(lengths[k]), lt, 0.0, models[k]) for k in idx]
    lp = O.lbfgs_params(m=10, ftol=1e-5, max_iter=10000, precondition=True)
    lib = O.load(omp=True)
    t1 = time.perf_counter()
    hs, rs = O.lbfgs_batch([np.full(lengths[k], ny / 2.0) for k in idx], po, tables, lp,
                           shifts=[shifts[k] for k in idx])
    sec = time.perf_counter() - t1
    e_err, it_diff, conv = [], [], 0
    for j, k in enumerate(idx):
        e_ref = O.objective(hs[j], po[j], tables[j])
        e_gpu = O.objective(h64[offs[k]:offs[k + 1]], po[j], tables[j])
        e_err.append(abs(e_gpu - e_ref) / abs(e_ref))
        it_diff.append(abs(res64[k].iterations - rs[j].iterations))
        conv += int(rs[j].converged)
    e_err = np.array(e_err)
    cpu = {"value": len(idx) / sec, "unit": "minima/s", "cores": int(lib.orc_num_threads()),
           "kind": "port", **omp_info,
           "sample": f"{len(idx)} lines of the batch (every {stride}th; lengths "
                     f"{sorted(set(lengths[k] for k in idx))}), oracle L-BFGS m=10 ftol 1e-5 "
                     f"preconditioned, OpenMP over lines, {sec:.1f} s"}
    parity = {"lines": len(idx), "oracle_converged": conv,
              "energy_rel_err_max": float(e_err.max()),
              "energy_rel_err_median": float(np.median(e_err)),
              "lines_within_1e-12": int((e_err <= 1e-12).sum()),
              "lines_within_1e-11": int((e_err <= 1e-11).sum()),
              "iterations_abs_diff_max": int(max(it_diff))}
    log(f"lbfgs cpu: {json.dumps(cpu)} parity {json.dumps(parity)}")
    return cpu, parity


def bench_langevin(log, nsteps=100000, with_oracle=True, omp_info=None):
    """Finite-temperature anneal: one line of L = 65536 sites on a 65536 x 4096
    FilteredNoise table (xi = 4, seed 1), linear elasticity, line tension 1,
    flat start at Ny/2, f = 0.5 f_c(est). Euler-Maruyama with dt = 0.01 for
    10^5 steps, temperature annealed linearly from 0.5 to 0.01 in the drivers'
    parameterisation (drift = T dt, diffusion = T sqrt(2 dt) for A = tau = 1,
    with the drivers' one-step lag, src/qew_runs.h:154-170), Philox normals
    keyed by (seed 2, step, site). FP64, FP32 (anchored offsets, FP32 table)
    and FP32 on the midpoint-contracted table run the same noise; the mean
    velocity, W^2 and the roughness exponent of the final line are compared.

    With `with_oracle`, the FP64 oracle runs the same anneal (all 10^5 steps,
    OpenMP over sites, one core fewer than the process may use) on a
    background thread while the caller goes on with other GPU work; the
    returned `finish()` joins it and adds the comparison with the oracle and
    the CPU baseline to the result."""
    import torch
    import qew_amd as q
    Lc, ny = 65536, 4096
    t0 = time.perf_counter()
    n64 = q.FilteredNoise(Lc, ny, float(Lc), float(ny), 1.0, 4.0, 4.0, 1, dtype=q.F64)
    n32 = n64.convert(q.F32)
    nmid = n64.convert(q.F32_MID)
    torch.cuda.synchronize()
    log(f"langevin: table generated in {time.perf_counter() - t0:.1f} s")
    f = 0.5 * FC_EST_LANGEVIN
    dt = 0.01
    total = nsteps * dt
    k = np.arange(nsteps)
    temp = np.where(k == 0, 0.5, 0.5 + (0.01 - 0.5) * ((k - 1) * dt / total))
    drift = temp * dt
    diff = temp * np.sqrt(2 * dt)
    p = q.Params(float(Lc), 1.0, f, q.Model.Linear)
    out = {}
    lines = {}
    for name, nz, dtype in (("f64", n64, q.F64), ("f32", n32, q.F32), ("f32_mid", nmid, q.F32)):
        b = q.Batch([Lc], [nz], [p])
        if dtype == q.F64:
            h = torch.full((Lc,), ny / 2.0, dtype=torch.float64, device="cuda:0")
            anchor = None
        else:
            h = torch.zeros(Lc, dtype=torch.float32, device="cuda:0")
            anchor = torch.full((1,), ny / 2.0, dtype=torch.float64, device="cuda:0")
        b.langevin(h.clone(), drift[:100], diff[:100], seed=2, step0=0, dtype=dtype, anchor=anchor)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        chunk = 10000
        for s in range(0, nsteps, chunk):
            b.langevin(h, drift[s:s + chunk], diff[s:s + chunk], seed=2, step0=s, dtype=dtype,
                       anchor=anchor)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        hh = h.double().cpu().numpy() + (ny / 2.0 if dtype == q.F32 else 0.0)
        lines[name] = hh
        # roughness exponent of the final line from its height-difference
        # ACF over 16 <= ell <= 4096 grid units
        from qew_amd import analysis
        out[name] = {"site_steps_per_s": nsteps * Lc / sec, "us_per_step": 1e6 * sec / nsteps,
                     "seconds": sec, "mean_velocity": float((hh.mean() - ny / 2.0) / total),
                     "W2": float(hh.var()),
                     "zeta": float(analysis.roughness_exponent(hh, float(Lc), 16.0, 4096.0))}
        log(f"langevin {name}: {json.dumps(out[name])}")
    out["f32_vs_f64_rel_err"] = {
        "mean_velocity": abs(out["f32"]["mean_velocity"] - out["f64"]["mean_velocity"])
        / abs(out["f64"]["mean_velocity"]),
        "W2": abs(out["f32"]["W2"] - out["f64"]["W2"]) / out["f64"]["W2"],
        "zeta_abs": abs(out["f32"]["zeta"] - out["f64"]["zeta"])}
    out["f32_mid_vs_f64_rel_err"] = {
        "mean_velocity": abs(out["f32_mid"]["mean_velocity"] - out["f64"]["mean_velocity"])
        / abs(out["f64"]["mean_velocity"]),
        "W2": abs(out["f32_mid"]["W2"] - out["f64"]["W2"]) / out["f64"]["W2"],
        "zeta_abs": abs(out["f32_mid"]["zeta"] - out["f64"]["zeta"])}
    out["workload"] = (f"L={Lc}, table {Lc}x{ny}, xi=4, linear, f=0.5*f_c(est)={f:g}, "
                       f"dt={dt}, {nsteps} steps, T 0.5 -> 0.01")
    log(f"langevin: {json.dumps(out)}")
    if not with_oracle:
        return out, lambda: None

    import threading
    O = oracle()
    coef = n64.coefficients()
    del n32, nmid, n64
    box = {}

    def anneal():
        lib = O.load(omp=True)
        lib.orc_set_num_threads(max(1, int(omp_info["OMP_NUM_THREADS"]) - 1))
        box["threads"] = int(lib.orc_num_threads())
        on = O.OracleNoise.table(coef, float(Lc), float(ny))
        po = O.params(float(Lc), 1.0, f, O.LINEAR)
        h_ref = np.full(Lc, ny / 2.0)
        t = time.perf_counter()
        O.langevin_run(h_ref, po, on, drift, diff, 2, 0, omp=True)
        box["seconds"] = time.perf_counter() - t
        box["h"] = h_ref

    th = threading.Thread(target=anneal, daemon=True)
    th.start()
    log("langevin: FP64 oracle anneal started on a background thread")

    def finish():
        th.join()
        h_ref = box["h"]
        v_ref = (h_ref.mean() - ny / 2.0) / total
        w2_ref = h_ref.var()
        out["vs_oracle"] = {
            name: {"mean_velocity_rel_err": abs(out[name]["mean_velocity"] - v_ref) / abs(v_ref),
                   "W2_rel_err": abs(out[name]["W2"] - w2_ref) / w2_ref,
                   "h_max_abs_err": float(np.max(np.abs(lines[name] - h_ref)))}
            for name in lines}
        out["vs_oracle"]["oracle"] = {"mean_velocity": float(v_ref), "W2": float(w2_ref)}
        out["cpu_baseline"] = {
            "value": nsteps * Lc / box["seconds"], "unit": "site-steps/s",
            "cores": box["threads"], "kind": "port", **omp_info,
            "sample": f"the whole anneal ({nsteps} steps of {Lc} sites), FP64 oracle with "
                      f"Philox normals, OpenMP over sites, {box['seconds']:.1f} s, run while the "
                      "GPU did the Rosso-Krauth workload"}
        log(f"langevin vs oracle: {json.dumps(out['vs_oracle'])} cpu {json.dumps(out['cpu_baseline'])}")

    return out, finish


def bench_rk(log, nlines=64, budget_s=60.0, precisions=("f64", "f32_mid"), k_way=4,
             with_cpu=True, omp_info=None, cpu_lines=2, state_path=None, defer_cpu=False):
    """Batched Rosso-Krauth depinning: nlines realizations of L = 16384 sites,
    each on its own 16384 x 1024 table (period M = 1024 in h, xi = 4, seeds
    1..nlines), linear elasticity, line tension 1, the depinning tool's knobs
    (dstep 0.5, max_advance Ly, runaway 2 Ly, root_tol 1e-10, sweep_tol 1e-8,
    max_sweeps 200000), f_c of every line to 1e-10 relative by
    qew_amd.depinning.fc_search (k_way forces per line per round) while
    `budget_s` lasts per precision:
    FP64 tables, and FP32 or midpoint-contracted FP32 tables (FP64 state,
    comparisons and bisection). Throughput: sweeps of the batch and of single
    lines per second, and the site-velocity evaluations actually performed.
    The CPU baseline runs the oracle (OpenMP over sites) through the same first
    two probes (relaxation at f = 0, ramp probe f = 0.125) on `cpu_lines`
    lines, whose heights must equal the GPU's bit for bit. With `state_path`
    (one precision), the search continues from the brackets and blocked
    configurations saved there by an earlier run and saves its own at the
    end, so that f_c to 1e-10 can be reached over several runs."""
    import torch
    import qew_amd as q
    from qew_amd.depinning import fc_search
    # probes near f_c sweep for minutes: rk_block reports its sweeps meanwhile
    os.environ.setdefault("QEW_RK_PROGRESS", "30")
    Lr, ny = 16384, 1024
    t0 = time.perf_counter()
    specs = [(Lr, ny, float(Lr), float(ny), 1.0, 4.0, 4.0, 1 + k) for k in range(nlines)]
    n64 = q.FilteredNoise.generate_batch(specs, dtype=q.F64)
    tables = {"f64": n64}
    if "f32" in precisions:
        tables["f32"] = [nz.convert(q.F32) for nz in n64]
    if "f32_mid" in precisions:
        # lines of n = L sites on L-wide tables of spacing 1: midpoint tables apply
        tables["f32_mid"] = [nz.convert(q.F32_MID) for nz in n64]
    torch.cuda.synchronize()
    log(f"rk: {nlines} tables generated in {time.perf_counter() - t0:.1f} s")
    rp = q.RkParams(0.5, float(ny), 1e-10, 1e-8, 2.0 * ny, 200000)
    out = {}
    brackets = {}
    first = {}

    def record(k, res, h, sec):
        first[k] = ([r.sweeps for r in res[:cpu_lines]], h[:cpu_lines].cpu().numpy(), sec,
                    sum(r.sweeps for r in res))

    for name in precisions:
        log(f"rk {name}:")
        state = None
        if state_path and os.path.exists(state_path):
            z = np.load(state_path)
            state = (z["lo"], z["hi"], z["h_lo"])
            log(f"rk {name}: continuing from {state_path}")
        def save(lo_, hi_, h_lo_):
            np.savez(state_path, lo=lo_, hi=hi_, h_lo=h_lo_.cpu().numpy())

        lo, hi, h_lo, st = fc_search(tables[name], Lr, ny / 2.0, rp, k_way=k_way,
                                     budget_s=budget_s, log=log,
                                     record=record if name == "f64" and state is None else None,
                                     state=state, checkpoint=save if state_path else None)
        brackets[name] = (lo, hi)
        width = (hi - lo) / np.maximum(np.abs(hi), 1e-300)
        sec = st.pop("seconds")
        out[name] = {"sweeps_per_s": st["batch_sweeps"] / sec,
                     "line_sweeps_per_s": st["line_sweeps"] / sec,
                     "velocity_evals_done_per_s": st["evals_done"] / sec,
                     "reference_equivalent_velocity_evals_per_s": st["evals_reference"] / sec,
                     "seconds": sec, **st,
                     "f_c_median": float(np.median(0.5 * (lo + hi))),
                     "f_c_min_max": [float(lo.min()), float(hi.max())],
                     "bracket_rel_width_max": float(width.max()),
                     "lines_at_1e-10": int((width <= 1e-10).sum())}
        log(f"rk {name}: {json.dumps({k: v for k, v in out[name].items() if k != 'round_log'})}")
    for name in precisions:
        if name == "f64" or "f64" not in brackets:
            continue
        lo, hi = brackets[name]
        lo64, hi64 = brackets["f64"]
        fc, fc64 = 0.5 * (lo + hi), 0.5 * (lo64 + hi64)
        out[f"{name}_vs_f64"] = {
            "f_c_rel_err_max": float(np.max(np.abs(fc - fc64) / fc64)),
            "identical_brackets": int(np.sum((lo == lo64) & (hi == hi64)))}
    out["workload"] = (f"{nlines} lines x L={Lr}, tables {Lr}x{ny} (period M={ny}), xi=4, "
                       f"linear, ramp + {k_way}-way bracketing of f_c to 1e-10, "
                       f"{budget_s:g} s budget per precision")
    log(f"rk: {json.dumps(out)}")
    if 0 in first and 1 in first:
        if with_cpu and not defer_cpu:
            out["cpu_baseline"], out["first_probes_vs_oracle"] = rk_oracle_sample(
                log, n64[:cpu_lines], Lr, ny, first, omp_info, nlines)
        elif with_cpu:
            # for the caller to run the oracle sample later (not serialised)
            out["first_probes"] = {"noises": n64[:cpu_lines], "Lr": Lr, "ny": ny,
                                   "first": first, "nlines": nlines}
    return out


def rk_compare(ref_path, path):
    """f_c brackets of two saved searches of the same batch (--rk-state files):
    the bracket widths reached, and f_c (bracket midpoints) of the second
    against the first."""
    a, b = np.load(ref_path), np.load(path)
    out = {}
    for name, z in (("ref", a), ("other", b)):
        lo, hi = z["lo"], z["hi"]
        width = (hi - lo) / np.maximum(np.abs(hi), 1e-300)
        out[name] = {"file": os.path.basename(ref_path if name == "ref" else path),
                     "lines": int(lo.size), "bracket_rel_width_max": float(width.max()),
                     "lines_at_1e-10": int((width <= 1e-10).sum()),
                     "f_c_min_max": [float(lo.min()), float(hi.max())],
                     "f_c_median": float(np.median(0.5 * (lo + hi)))}
    fa, fb = 0.5 * (a["lo"] + a["hi"]), 0.5 * (b["lo"] + b["hi"])
    rel = np.abs(fb - fa) / np.abs(fa)
    out["other_vs_ref"] = {"f_c_rel_err_max": float(rel.max()),
                           "f_c_rel_err_median": float(np.median(rel)),
                           "identical_brackets": int(np.sum((a["lo"] == b["lo"]) &
                                                            (a["hi"] == b["hi"]))),
                           "brackets_overlap": int(np.sum((np.maximum(a["lo"], b["lo"]) <=
                                                           np.minimum(a["hi"], b["hi"]))))}
    return out


def rk_oracle_sample(log, noises, Lr, ny, first, omp_info, nlines):
    """The oracle's rk_block (OpenMP over the sites of a colour) through the
    relaxation at f = 0 and the ramp probe f = 0.125 on the first lines of the
    batch: line sweeps per second against the GPU's for the same two probes on
    all lines, and the heights and sweep counts of those lines bit for bit."""
    O = oracle()
    lib = O.load(omp=True)
    rpo = O.rk_params(dstep=0.5, max_advance=float(ny), root_tol=1e-10, sweep_tol=1e-8,
                      runaway=2.0 * ny, max_sweeps=200000)
    sweeps, same, sec = 0, True, 0.0
    for j, nz in enumerate(noises):
        on = O.OracleNoise.table(nz.coefficients(), float(Lr), float(ny))
        h = np.full(Lr, ny / 2.0)
        for k, f in ((0, 0.0), (1, 0.125)):
            t = time.perf_counter()
            r = O.rk_block(h, O.params(float(Lr), 1.0, f, O.LINEAR), on, rpo, omp=True)
            sec += time.perf_counter() - t
            sweeps += r.sweeps
            same = same and r.sweeps == first[k][0][j] and np.array_equal(h, first[k][1][j])
    gpu_sec = first[0][2] + first[1][2]
    gpu_line_sweeps = first[0][3] + first[1][3]
    cpu = {"value": sweeps / sec, "unit": "line sweeps/s (16384 sites)",
           "cores": int(lib.orc_num_threads()), "kind": "port", **omp_info,
           "sample": f"relaxation at f=0 and the ramp probe f=0.125 on {len(noises)} lines "
                     f"({sweeps} line sweeps, {sec:.1f} s)",
           "gpu_same_probes": {"value": gpu_line_sweeps / gpu_sec,
                               "sample": f"the same two probes on all {nlines} lines, "
                                         f"{gpu_sec:.2f} s"}}
    parity = {"lines": len(noises), "heights_and_sweeps_bit_identical": bool(same)}
    log(f"rk cpu: {json.dumps(cpu)} {json.dumps(parity)}")
    return cpu, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--evals", type=int, default=EVALS)
    ap.add_argument("--shift", type=int, default=SHIFT,
                    help="cells between the walks of consecutive evaluations (mod Ny)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lbfgs", action="store_true", help="skip the batched L-BFGS workload")
    ap.add_argument("--no-langevin", action="store_true", help="skip the Langevin anneal")
    ap.add_argument("--langevin-steps", type=int, default=100000,
                    help="anneal length (the workload is 10^5 steps; fewer only for profiling)")
    ap.add_argument("--no-rk", action="store_true", help="skip the batched Rosso-Krauth workload")
    ap.add_argument("--rk-budget", type=float, default=20.0,
                    help="seconds of f_c search per precision in the Rosso-Krauth workload "
                         "(checked between rounds; f_c to 1e-10 takes about 15 rounds)")
    ap.add_argument("--rk-only", action="store_true",
                    help="run only the Rosso-Krauth workload and print its JSON object")
    ap.add_argument("--rk-state", default=None,
                    help="with --rk-only and one precision: continue the f_c search from this "
                         ".npz file if it exists, and save the search state there after the "
                         "ramp and after every round")
    ap.add_argument("--rk-compare", nargs=2, metavar=("REF_NPZ", "NPZ"), default=None,
                    help="compare two saved f_c search states (brackets per line; the first "
                         "is the reference precision) and print the JSON summary (CPU only)")
    ap.add_argument("--rk-precisions", default="f64,f32_mid",
                    help="table types of the Rosso-Krauth workload (f64, f32, f32_mid)")
    ap.add_argument("--rk-kway", type=int, default=4,
                    help="forces probed per line per round of the f_c search")
    ap.add_argument("--lbfgs-lines", type=int, default=512)
    ap.add_argument("--lbfgs-only", action="store_true",
                    help="run only the batched L-BFGS workload and print its JSON object")
    ap.add_argument("--only", choices=["both", "f32", "f64"], default="both")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-table evaluation leg")
    ap.add_argument("--no-mid", action="store_true",
                    help="skip the midpoint-contracted FP32 table evaluation leg")
    ap.add_argument("--count-lines", action="store_true",
                    help="only count the 128-byte table lines one evaluation of the walk touches "
                         "under candidate table layouts (CPU) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (reference point for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE scaling) and exit")
    args = ap.parse_args()
    if args.rk_compare:
        print(json.dumps(rk_compare(*args.rk_compare)), flush=True)
        return
    if args.count_lines:
        print(json.dumps(layout_report(walk(), L, L, NY)), flush=True)
        return
    omp_info = omp_setup()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the engine runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    import qew_amd as q
    unpin()

    dev = torch.device("cuda:0")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.lbfgs_only:
        out = bench_lbfgs(log, nlines=args.lbfgs_lines, with_cpu=not args.no_cpu,
                          omp_info=omp_info)
        print(json.dumps(out), flush=True)
        return
    if args.rk_only:
        out = bench_rk(log, budget_s=args.rk_budget,
                       precisions=tuple(args.rk_precisions.split(",")), k_way=args.rk_kway,
                       with_cpu=not args.no_cpu, omp_info=omp_info, state_path=args.rk_state)
        print(json.dumps(out), flush=True)
        return
    h0 = walk()
    B = args.evals
    shifts = torch.tensor(((np.arange(B) * args.shift) % NY).astype(np.float64), device=dev)
    u_sum = distinct_coefficients(h0, L, L, NY)
    log(f"sum_r U_r = {u_sum} ({u_sum / L:.3f} per row)")
    u_mid = distinct_mid_values(h0, L, L, NY)

    results = {}
    params = q.Params(float(L), 1.0, 0.0, q.Model.Arclength)

    def run(table):
        # table storage "f64" (FP64 line state), "f32", "bf16" or "f32_mid"
        # (FP32 line state; f32_mid: the FP64 table contracted at the edge
        # midpoints' x-fraction 1/2)
        t_gen = time.perf_counter()
        dtype = q.F64 if table == "f64" else q.F32
        if table == "f32_mid":
            base = results.get("table64") or q.FilteredNoise(L, NY, float(L), float(NY), 1.0, XI,
                                                              XI, 1, dtype=q.F64)
            noise = base.convert(q.F32_MID)
            del base
        else:
            noise = q.FilteredNoise(L, NY, float(L), float(NY), 1.0, XI, XI, 1, dtype=dtype)
        if table == "bf16":
            noise = noise.convert(q.BF16)
        torch.cuda.synchronize()
        log(f"table {table} generated in {time.perf_counter() - t_gen:.1f} s")
        batch = q.Batch([L] * B, [noise] * B, [params] * B)
        h0d = torch.tensor(h0, dtype=torch.float64, device=dev)
        if dtype == q.F64:
            h = (h0d[None, :] + shifts[:, None]).reshape(-1).contiguous()
            anchor = None
            grad = torch.empty_like(h)
        else:
            off = (h0d - NY / 2).to(torch.float32)
            h = off[None, :].expand(B, L).contiguous().reshape(-1)
            anchor = (NY / 2 + shifts).contiguous()
            grad = torch.empty_like(h)
        energy = torch.empty(B, dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            batch.eval(h, grad, energy, dtype=dtype, anchor=anchor)
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            starts[k].record()
            batch.eval(h, grad, energy, dtype=dtype, anchor=anchor)
            ends[k].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
        s_c = {"f64": 8, "f32": 4, "bf16": 2, "f32_mid": 4}[table]  # table bytes per value
        s_h = 8 if dtype == q.F64 else 4               # line state and gradient bytes
        values = u_mid if table == "f32_mid" else u_sum
        bytes_eval = s_c * values + L * 2 * s_h
        res = {
            "evals_per_s": B * args.steps / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "kernel_ms": kms,
            "bytes_per_eval": bytes_eval,
            "achieved_GBs": bytes_eval * B / (kms * 1e-3) / 1e9,
        }
        first = (h[:L].clone(), grad[:L].clone(), float(energy[0]))
        if table == "f64":
            results["table64"] = noise
        del batch, h, grad, energy
        if table != "f64":
            del noise
        torch.cuda.empty_cache()
        return res, first

    r32 = r64 = r16 = None
    if args.only in ("both", "f64"):
        r64, first64 = run("f64")
        log(f"f64: {json.dumps(r64)}")
    if args.only in ("both", "f32"):
        r32, first32 = run("f32")
        log(f"f32: {json.dumps(r32)}")
    if args.only == "both" and not args.no_bf16:
        r16, first16 = run("bf16")
        log(f"bf16: {json.dumps(r16)}")
    rmid = None
    if args.only == "both" and not args.no_mid:
        rmid, firstmid = run("f32_mid")
        log(f"f32_mid: {json.dumps(rmid)}")

    def error_vs_f64(first):
        g64 = first64[1].cpu().numpy()
        g = first[1].cpu().numpy().astype(np.float64)
        return {"grad_max_abs_err_over_max_grad": float(np.max(np.abs(g - g64)) / np.max(np.abs(g64))),
                "energy_rel_err": abs(first[2] - first64[2]) / abs(first64[2])}

    err = error_vs_f64(first32) if (r32 and r64) else None

    cpu = None
    if not args.no_cpu and "table64" in results:
        coef = results["table64"].coefficients()
        del results["table64"]
        cpu = cpu_baseline(coef, h0)
        cpu.update(omp_info)
        log(f"cpu: {json.dumps(cpu)}")

    head = r32 if r32 else r64
    head_dtype = "f32" if r32 else "f64"
    traffic = traffic_from_profiles()
    traffic_launch = None
    if traffic and head_dtype in traffic.get("detail", {}):
        d = traffic["detail"][head_dtype]
        # HBM bytes of one launch of B evaluations (rocprof FETCH_SIZE + WRITE_SIZE per evaluation)
        traffic_launch = (d["fetch_bytes_per_eval"] + d["write_bytes_per_eval"]) * B
    line = {
        "metric": "grad evals/s + % HBM peak at L=65536; batched L-BFGS minima/s, FP32 vs FP64",
        "value": head["evals_per_s"],
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "kernel_ms": head["kernel_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head_dtype,
        "data": "synthetic (seeded FilteredNoise table, seeded random-walk line)",
        "config": {"workload": f"fused gradient+energy, L={L}, table {L}x{NY}, xi={XI:g}, arclength, "
                               f"{B} evaluations per step (one batched launch)",
                   "L": L, "noise_grid": [L, NY], "xi": XI, "evals_per_step": B},
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": traffic_launch},
        "cpu_baseline": cpu,
    }
    if r64:
        line["f64"] = {"value": r64["evals_per_s"], "ms_per_step": r64["ms_per_step"],
                       "kernel_ms": r64["kernel_ms"],
                       "roofline_achieved_GBs": r64["achieved_GBs"],
                       "roofline_frac": r64["achieved_GBs"] / HBM_PEAK_GBS,
                       "bytes_per_eval": r64["bytes_per_eval"]}
    if r32:
        line["f32_bytes_per_eval"] = r32["bytes_per_eval"]
    if err:
        line["fp32_vs_fp64_error"] = err
    if r16:
        # BF16 coefficient storage, FP32 line state and arithmetic (optional mode)
        line["bf16_table"] = {"value": r16["evals_per_s"], "ms_per_step": r16["ms_per_step"],
                              "kernel_ms": r16["kernel_ms"],
                              "bytes_per_eval": r16["bytes_per_eval"],
                              "roofline_achieved_GBs": r16["achieved_GBs"],
                              "roofline_frac": r16["achieved_GBs"] / HBM_PEAK_GBS,
                              "error_vs_f64": error_vs_f64(first16) if r64 else None}
    if rmid:
        # midpoint-contracted FP32 table: 4 values per edge instead of the
        # 4 x 4 window (every edge of this line lies at x-fraction 1/2)
        line["f32_mid_table"] = {"value": rmid["evals_per_s"], "ms_per_step": rmid["ms_per_step"],
                                 "kernel_ms": rmid["kernel_ms"],
                                 "bytes_per_eval": rmid["bytes_per_eval"],
                                 "roofline_achieved_GBs": rmid["achieved_GBs"],
                                 "roofline_frac": rmid["achieved_GBs"] / HBM_PEAK_GBS,
                                 "error_vs_f64": error_vs_f64(firstmid) if r64 else None}
    with_cpu = not args.no_cpu
    if not args.no_lbfgs:
        torch.cuda.empty_cache()
        line["batched_lbfgs"] = bench_lbfgs(log, nlines=args.lbfgs_lines, with_cpu=with_cpu,
                                            omp_info=omp_info)
    finish_langevin = lambda: None
    if not args.no_langevin:
        torch.cuda.empty_cache()
        # the FP64 oracle anneal runs on host threads during the next leg
        line["langevin"], finish_langevin = bench_langevin(
            log, nsteps=args.langevin_steps, with_oracle=with_cpu, omp_info=omp_info)
    if not args.no_rk:
        torch.cuda.empty_cache()
        line["rosso_krauth"] = bench_rk(log, budget_s=args.rk_budget,
                                        precisions=tuple(args.rk_precisions.split(",")),
                                        k_way=args.rk_kway, with_cpu=with_cpu, defer_cpu=True,
                                        omp_info=omp_info)
    finish_langevin()
    if not args.no_rk and with_cpu:
        # the oracle's Rosso-Krauth sample after the anneal has freed the cores
        rk = line["rosso_krauth"]
        if "first_probes" in rk:
            fp = rk.pop("first_probes")
            rk["cpu_baseline"], rk["first_probes_vs_oracle"] = rk_oracle_sample(
                log, fp["noises"], fp["Lr"], fp["ny"], fp["first"], omp_info, fp["nlines"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
