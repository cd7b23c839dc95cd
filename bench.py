"""Benchmark of the qEW gradient/objective hot path on one MI355X.

Workload (the gradient/objective throughput configuration): a line of
L = 65536 sites on a 65536 x 8192 FilteredNoise table (xi = 8 grid units,
unit amplitude, seed 1, grid units dx = dy = 1), arclength elasticity,
line tension 1, no driving force. The line is a seeded random walk with
unit-variance Gaussian steps started at Nh/2. One step = 10^4 fused
gradient + energy evaluations issued as one batched launch; evaluation r sees
the walk shifted by 8 r cells (mod Nh), so the evaluations in flight read
disjoint table rows and the 1024-shift cycle spans the whole table (no L2/MALL
reuse across evaluations). Line states and gradients of all 10^4 evaluations
are distinct buffers in HBM.

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
SHIFT = 8
NSHIFT = NY // SHIFT
# Rosso-Krauth estimate of the depinning force of the Langevin workload's
# disorder (65536 x 4096 table, xi = 4, seed 1, linear, line tension 1, FP32
# table, the depinning tool's knobs): blocked at 0.5078, not blocked at
# 0.5156 (a 2*10^5-sweep timeout); ramp and bisection as in bench_rk.
FC_EST_LANGEVIN = 0.512

ROOT = os.path.dirname(os.path.abspath(__file__))


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
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    lib = O.load(omp=True)
    nz = O.OracleNoise.table(coef_rowmajor, float(L), float(NY))
    p = O.params(float(L), 1.0, 0.0, O.ARCLENGTH)
    t0 = time.perf_counter()
    k = 0
    while True:
        h = h0 + SHIFT * (k % NSHIFT)
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


def bench_lbfgs(log, nlines=512, plain_sample=8):
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
    n32 = [nz.convert(q.F32) for nz in n64]
    # every line has n = L sites on an L-wide table of spacing 1: its edge
    # midpoints lie at x-fraction 1/2, so midpoint-contracted tables apply
    nmid = [nz.convert(q.F32_MID) for nz in n64]
    n64mid = [nz.convert(q.F64_MID) for nz in n64]
    torch.cuda.synchronize()
    log(f"lbfgs: {nlines} tables generated in {time.perf_counter() - t0:.1f} s "
        f"(sum L = {sum(lengths)})")
    lt = 4.0 ** 1.5
    params = [q.Params(float(L_), lt, 0.0, q.Model(m)) for L_, m in zip(lengths, models)]
    shifts = [0.5 / xi ** 2 for xi in xis]
    h0 = torch.full((sum(lengths),), ny / 2.0, dtype=torch.float64, device="cuda:0")
    out = {}
    finals, gmax, minima = {}, {}, {}
    offs = np.concatenate([[0], np.cumsum(lengths)])
    # ftol 1e-5: above the Armijo roundoff floor of the longest lines (L=16384,
    # |E| ~ 1e4), where the reference algorithm converges in ~100-300 iterations
    # FP32: the line offsets' resolution floors the FP32 gradient near 1e-4, so
    # FP32 stops there and is compared by its energy against the FP64 minimum
    for name, noises, dtype, ftol in (("f64", n64, q.F64, 1e-5), ("f32", n32, q.F32, 1e-4),
                                      ("f32_mid", nmid, q.F32, 1e-4),
                                      ("f64_mid", n64mid, q.F64, 1e-5)):
        b = q.Batch(lengths, noises, params)
        b.set_precond_shifts(shifts)
        lp = q.LbfgsParams(m=10, ftol=ftol, max_iter=10000, precondition=True)
        h = h0.clone()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = b.lbfgs(h, lp, dtype=dtype)
        dt = time.perf_counter() - t1
        # energies and gradients of the minima, evaluated in FP64 on the FP64 tables
        e, g = q.Batch(lengths, n64, params).eval(h)
        finals[name] = e.cpu().numpy()
        minima[name] = h.cpu().numpy()
        ga = g.abs().cpu().numpy()
        gmax[name] = np.array([ga[o0:o1].max() for o0, o1 in zip(offs[:-1], offs[1:])])
        its = [r.iterations for r in res]
        out[name] = {"minima_per_s": nlines / dt, "seconds": dt, "ftol": ftol,
                     "converged": int(sum(r.converged for r in res)),
                     "curvature_skips": int(sum(r.n_curvature_skips for r in res)),
                     "ls_restarts": int(sum(r.n_ls_restarts for r in res)),
                     "slowest_line": int(np.argmax(its)),
                     "iterations_mean": float(np.mean(its)), "iterations_max": int(max(its))}
        log(f"lbfgs {name}: {json.dumps(out[name])}")
        del b, h
    rel_err = np.abs(finals["f32"] - finals["f64"]) / np.abs(finals["f64"])
    out["f32_vs_f64_energy_rel_err"] = {"max": float(rel_err.max()),
                                        "median": float(np.median(rel_err)),
                                        "lines_within_1e-6": int((rel_err <= 1e-6).sum())}
    rel_m64 = np.abs(finals["f64_mid"] - finals["f64"]) / np.abs(finals["f64"])
    out["f64_mid_vs_f64_energy_rel_err"] = {"max": float(rel_m64.max()),
                                            "median": float(np.median(rel_m64)),
                                            "lines_within_1e-6": int((rel_m64 <= 1e-6).sum())}
    rel_mid = np.abs(finals["f32_mid"] - finals["f64"]) / np.abs(finals["f64"])
    out["f32_mid_vs_f64_energy_rel_err"] = {"max": float(rel_mid.max()),
                                            "median": float(np.median(rel_mid)),
                                            "lines_within_1e-6": int((rel_mid <= 1e-6).sum())}
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
    out["workload"] = (f"{nlines} lines, L from {{1024..16384}} (sum {sum(lengths)}), "
                       f"tables L x {ny}, xi in {{2,4,8}}, m=10, FFT preconditioner")
    log(f"lbfgs: {json.dumps(out)}")
    return out


def bench_langevin(log, nsteps=100000):
    """Finite-temperature anneal: one line of L = 65536 sites on a 65536 x 4096
    FilteredNoise table (xi = 4, seed 1), linear elasticity, line tension 1,
    flat start at Ny/2, f = 0.5 f_c(est). Euler-Maruyama with dt = 0.01 for
    10^5 steps, temperature annealed linearly from 0.5 to 0.01 in the drivers'
    parameterisation (drift = T dt, diffusion = T sqrt(2 dt) for A = tau = 1,
    with the drivers' one-step lag, src/qew_runs.h:154-170), Philox normals
    keyed by (seed 2, step, site). FP64, FP32 (anchored offsets, FP32 table)
    and FP32 on the midpoint-contracted table run the same noise; the mean
    velocity, W^2 and the roughness exponent of the final line are compared."""
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
    return out


def bench_rk(log, nlines=64, budget_s=20.0):
    """Batched Rosso-Krauth depinning: nlines realizations of L = 16384 sites,
    each on its own 16384 x 1024 table (period M = 1024 in h, xi = 4, seeds
    1..nlines), linear elasticity, line tension 1, the depinning tool's knobs
    (dstep 0.5, max_advance Ly, runaway 2 Ly, root_tol 1e-10, sweep_tol 1e-8,
    max_sweeps 200000). Relax at f = 0, ramp f = 0.125 * 2^k per line until it
    runs away, then bisect every line's bracket, each probe warm-started from
    the line's highest blocked configuration (no-passing). Probes are issued
    while `budget_s` lasts; near f_c a probe takes 10^4..2*10^5 sweeps, so the
    bracket after the budget is reported, not f_c to 1e-10. FP32 tables (FP64
    state, comparisons and bisection) and FP64 tables run the same probes."""
    import torch
    import qew_amd as q
    Lr, ny = 16384, 1024
    t0 = time.perf_counter()
    specs = [(Lr, ny, float(Lr), float(ny), 1.0, 4.0, 4.0, 1 + k) for k in range(nlines)]
    n64 = q.FilteredNoise.generate_batch(specs, dtype=q.F64)
    n32 = [nz.convert(q.F32) for nz in n64]
    # lines of n = L sites on L-wide tables of spacing 1: midpoint tables apply
    nmid = [nz.convert(q.F32_MID) for nz in n64]
    torch.cuda.synchronize()
    log(f"rk: {nlines} tables generated in {time.perf_counter() - t0:.1f} s")
    params = [q.Params(float(Lr), 1.0, 0.0, q.Model.Linear)] * nlines
    rp = q.RkParams(0.5, float(ny), 1e-10, 1e-8, 2.0 * ny, 200000)
    out = {}
    brackets = {}
    for name, noises in (("f32", n32), ("f64", n64), ("f32_mid", nmid)):
        b = q.Batch([Lr] * nlines, noises, params)
        h_lo = torch.full((nlines * Lr,), ny / 2.0, dtype=torch.float64, device="cuda:0")
        lo = np.zeros(nlines)
        hi = np.full(nlines, np.inf)
        evals = 0
        sweeps = 0
        probes = 0
        f_ramp = 0.125
        t1 = time.perf_counter()
        forces = np.zeros(nlines)
        while True:
            b.set_forces(forces)
            h = h_lo.clone()
            res = b.rk_block(h, rp)
            probes += 1
            evals += sum(r.n_velocity_evals for r in res)
            sweeps += max(r.sweeps for r in res)
            blocked = np.array([r.blocked for r in res])
            if probes == 1:
                assert blocked.all(), "a line does not block at f = 0"
                h_lo = h
            else:
                m = torch.tensor(np.repeat(blocked & todo, Lr), device="cuda:0")
                h_lo = torch.where(m, h, h_lo)
                lo = np.where(blocked & todo, forces, lo)
                hi = np.where(~blocked & todo, forces, hi)
            todo = np.isinf(hi) | (hi - lo > 1e-10 * np.maximum(np.abs(lo), np.abs(hi)))
            if not todo.any() or time.perf_counter() - t1 > budget_s:
                break
            forces = np.where(np.isinf(hi), np.maximum(2 * lo, f_ramp), 0.5 * (lo + hi))
            forces = np.where(todo, forces, lo)  # finished lines: one blocked sweep
            f_ramp *= 2
        sec = time.perf_counter() - t1
        brackets[name] = (lo.copy(), hi.copy())
        width = (hi - lo) / np.maximum(hi, 1e-300)
        out[name] = {"velocity_evals_per_s": evals / sec, "sweeps_per_s": sweeps / sec,
                     "seconds": sec, "probes": probes, "sweeps": sweeps,
                     "velocity_evals": int(evals),
                     "f_c_bracket_median": [float(np.median(lo)), float(np.median(hi))],
                     "bracket_rel_width_max": float(width.max())}
        log(f"rk {name}: {json.dumps(out[name])}")
        del b, h, h_lo
    same = (np.array_equal(brackets["f32"][0], brackets["f64"][0])
            and np.array_equal(brackets["f32"][1], brackets["f64"][1]))
    out["f32_f64_same_brackets"] = bool(same)
    out["f32_mid_f64_same_brackets"] = bool(
        np.array_equal(brackets["f32_mid"][0], brackets["f64"][0])
        and np.array_equal(brackets["f32_mid"][1], brackets["f64"][1]))
    out["workload"] = (f"{nlines} lines x L={Lr}, tables {Lr}x{ny} (period M={ny}), xi=4, "
                       f"linear, f_c ramp + bisection, {budget_s:g} s budget per precision")
    out["velocity_evals_note"] = ("evaluations the reference algorithm makes (sites whose "
                                  "inputs did not change are skipped on the GPU but counted)")
    log(f"rk: {json.dumps(out)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--evals", type=int, default=EVALS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lbfgs", action="store_true", help="skip the batched L-BFGS workload")
    ap.add_argument("--no-langevin", action="store_true", help="skip the Langevin anneal")
    ap.add_argument("--langevin-steps", type=int, default=100000,
                    help="anneal length (the workload is 10^5 steps; fewer only for profiling)")
    ap.add_argument("--no-rk", action="store_true", help="skip the batched Rosso-Krauth workload")
    ap.add_argument("--rk-budget", type=float, default=20.0,
                    help="seconds of f_c bisection per precision in the Rosso-Krauth workload")
    ap.add_argument("--lbfgs-lines", type=int, default=512)
    ap.add_argument("--only", choices=["both", "f32", "f64"], default="both")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-table evaluation leg")
    ap.add_argument("--no-mid", action="store_true",
                    help="skip the midpoint-contracted FP32 table evaluation leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (reference point for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE scaling) and exit")
    args = ap.parse_args()
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

    dev = torch.device("cuda:0")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    h0 = walk()
    B = args.evals
    shifts = torch.tensor((np.arange(B) % NSHIFT) * SHIFT, dtype=torch.float64, device=dev)
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
                              "bytes_per_eval": r16["bytes_per_eval"],
                              "roofline_achieved_GBs": r16["achieved_GBs"],
                              "roofline_frac": r16["achieved_GBs"] / HBM_PEAK_GBS,
                              "error_vs_f64": error_vs_f64(first16) if r64 else None}
    if rmid:
        # midpoint-contracted FP32 table: 4 values per edge instead of the
        # 4 x 4 window (every edge of this line lies at x-fraction 1/2)
        line["f32_mid_table"] = {"value": rmid["evals_per_s"], "ms_per_step": rmid["ms_per_step"],
                                 "bytes_per_eval": rmid["bytes_per_eval"],
                                 "roofline_achieved_GBs": rmid["achieved_GBs"],
                                 "roofline_frac": rmid["achieved_GBs"] / HBM_PEAK_GBS,
                                 "error_vs_f64": error_vs_f64(firstmid) if r64 else None}
    if not args.no_lbfgs:
        torch.cuda.empty_cache()
        line["batched_lbfgs"] = bench_lbfgs(log, nlines=args.lbfgs_lines)
    if not args.no_langevin:
        torch.cuda.empty_cache()
        line["langevin"] = bench_langevin(log, nsteps=args.langevin_steps)
    if not args.no_rk:
        torch.cuda.empty_cache()
        line["rosso_krauth"] = bench_rk(log, budget_s=args.rk_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
