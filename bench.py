"""Benchmark: Davidson time-to-1e-6 Eh of the CI hot path on one MI355X, FP64 vs mixed.

Workload (default): electronic full CI with synthetic seeded integrals, 16 orbitals, 8 alpha +
8 beta electrons (12870^2 = 165,636,900 determinants), 4 lowest roots, Davidson residual
<= 1e-6 (search space: 16 start vectors, 48 at most, start perturbation 1e-2), determinant set,
integrals, tables, subspace and every vector resident in HBM.

A step is one Davidson iteration of the plain FP64 solve (the sigma of the new block of search
directions, Rayleigh-Ritz, Ritz vectors, residuals, corrections and their orthonormalisation;
one host synchronisation).  The solve runs an iteration at a time: the first W iterations are
the warm-up, the next K are timed (ms_per_step), and the solve then continues to convergence.
value = 1 / (FP64 time-to-1e-6: start vectors to converged roots) in solves per second.  The
mixed-precision solve of the same problem (FP32 sigma bulk, FP64 correction, subspace and final
residuals) is timed beside it ("mixed"), with its per-root difference from the FP64 energies.

roofline: the dominant kernel is the link GEMM k_lg_u (FP64 MFMA), every off-diagonal term of
the sigma.  Its launch durations come from HIP events on the CI stream during one extra FP64
sigma of 4 vectors; achieved = the algorithmic flops of that sigma, F = 2 * k * N * (conn + 1)
with conn = 2 s + 2 d + s^2 = 5792 connected determinants per determinant (s = 64 singles,
d = 784 same-spin doubles per string) and k = 4, divided by the summed k_lg_u time; the GEMM's
own flop count (its M x K x N as posed) is reported beside it.  The sigma's HBM-bound kernel, the
gather of U into sigma (k_lg_gather), is reported in sigma_kernels against the 8 TB/s HBM peak on
its algorithmic bytes, k * N * w * (2 L + 3): U read once per pass (L = 64 links per string),
sigma written in both passes and read back in the accumulating one (w = 8 B FP64, 4 B mixed).
cpu_baseline: the oracle's restated reference sigma on strided determinant samples of the same
space and vectors, per-row cost scaled to the full sigma and multiplied by the number of sigma
vectors of the GPU FP64 solve (an extrapolation, labelled so in "sample").

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
    python bench.py --workload ep48|sigma|ao2mo|serve|h2o|rdm   (the other measured workloads)
"""
import argparse
import math
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(prob, sigma_calls, nvec, threads):
    """Oracle sigma on rows i = 0, s, 2s, ... (two strides fit the fixed per-call overhead)."""
    import numpy as np

    from oracle import refcpu

    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(0).standard_normal((oc.N, nvec))
    s_big, s_small = 256, 32
    t0 = time.perf_counter()
    oc.sigma_sample(C, s_big)
    t_big = time.perf_counter() - t0
    t0 = time.perf_counter()
    oc.sigma_sample(C, s_small)
    t_small = time.perf_counter() - t0
    rows_big, rows_small = -(-oc.N // s_big), -(-oc.N // s_small)
    per_row = max((t_small - t_big) / (rows_small - rows_big), 1e-12)
    overhead = max(t_big - per_row * rows_big, 0.0)
    t_sigma = overhead + per_row * oc.N
    t_solve = t_sigma * sigma_calls
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle symmetric-scatter sigma on rows i = 0 mod {s_small} ({rows_small} of {oc.N} "
                   f"determinants, k={nvec}) and mod {s_big}; per-row cost {per_row * 1e6:.1f} us wall on "
                   f"{threads} threads -> full sigma {t_sigma:.2f} s x {sigma_calls} sigma calls of the GPU solve"),
        "sigma_s": t_sigma,
    }


PMC_SUMMARY = {"tensor_gemm_f64": "k_pair_gemm_f64_ep_fci_48e_48p_pmc.md",
               "sigma_pass_nc3_f64": "k_sigma_pass_nc3_f64_ep_fci_48e_48p_pmc.md",
               "sigma_pass_nc1_f64": "k_sigma_pass_nc1_f64_ep_fci_48e_48p_pmc.md",
               "rdm_group": "k_rdm_group_fci_16o_8a8b.md",
               "lg_u_f64": "k_lg_u_f64_fci_16o_8a8b_pmc.md"}


def traffic_from_profiles(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", PMC_SUMMARY[kernel])
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def bench_ao2mo(args):
    """AO->MO transform, n_ao = n_mo = 320, 1280 Cholesky vectors (one JSON line)."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd.ci import Context
    from polyquant_amd.integral import BF16X3, FP64, cholesky_eri8_device, transform_device

    assert torch.cuda.is_available(), "no GPU visible"
    n, nvec = args.ao_n, 1280
    npr = n * (n + 1) // 2
    t0 = time.perf_counter()
    L, C = wl.config4(n=n, nvec=nvec)
    log(f"[ao2mo] n={n} npair={npr} inputs generated in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda")
    # 8-fold packed AO integrals: eri8[idx2(P, Q)] = sum_k L[k, P] L[k, Q]
    eri8, Lg = cholesky_eri8_device(L)
    del L
    Cg = torch.from_numpy(np.ascontiguousarray(C.ravel(order="F"))).to(dev)
    out64 = torch.empty(npr * npr, dtype=torch.float64, device=dev)
    out3 = torch.empty(npr * npr, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"[ao2mo] AO integrals built on the GPU ({eri8.numel() * 8 / 1e9:.2f} GB 8-fold packed)")
    ctx = Context(0, hbm_budget=int(args.ao_budget_gb * (1 << 30)))
    f_alg = 8.0 * n ** 3 * npr
    res = {}
    for mode, name, out, peak in ((FP64, "fp64", out64, 78.6), (BF16X3, "bf16x3", out3, 2500.0 / 3)):
        gemm = "ao2mo_gemm_f64" if mode == FP64 else "ao2mo_gemm_bf16x3"
        times = []
        for rep in range(args.ao_reps):
            ctx.profile(True)
            ctx.profile_read(gemm)
            sec = transform_device(ctx, n, eri8.data_ptr(), Cg.data_ptr(), out.data_ptr(), mode)
            ctx.profile(False)
            launches, gemm_ms, units = ctx.profile_read(gemm)
            times.append((sec, gemm_ms, launches, units))
            log(f"[ao2mo] {name} rep {rep}: {sec * 1e3:.1f} ms total, GEMM launches {launches} {gemm_ms:.1f} ms")
        sec, gemm_ms, launches, units = min(times)
        res[name] = {"seconds": round(sec, 5), "tflops": round(f_alg / sec / 1e12, 3),
                     "frac_of_peak": round(f_alg / sec / 1e12 / peak, 4), "peak_tflops": round(peak, 1),
                     "gemm_ms": round(gemm_ms, 3), "gemm_launches": launches,
                     "gemm_tflops": round(units / max(gemm_ms * 1e-3, 1e-30) / 1e12, 3)}
    torch.cuda.synchronize()
    # BF16x3 vs FP64 on every element
    dmax, amax = 0.0, 0.0
    step = 1 << 28
    for i in range(0, npr * npr, step):
        a, b = out64[i:i + step], out3[i:i + step]
        dmax = max(dmax, float((a - b).abs().max()))
        amax = max(amax, float(a.abs().max()))
    # FP64 result against the closed form on sampled columns: (pq|rs) = sum_k (C^T L_k C)_pq (C^T L_k C)_rs
    P = torch.from_numpy(wl.pair_index_matrix(n).astype(np.int64)).to(dev)
    Cm = Cg.reshape(n, n).T.contiguous()  # row-major C[mu][i]
    ii, jj = np.tril_indices(n)
    flat = torch.from_numpy((ii * n + jj).astype(np.int64)).to(dev)
    Lmo = torch.empty(nvec, npr, dtype=torch.float64, device=dev)
    for k0 in range(0, nvec, 128):
        Lk = Lg[k0:k0 + 128][:, P.reshape(-1)].reshape(-1, n, n)
        Mk = torch.matmul(torch.matmul(Cm.T.contiguous(), Lk), Cm)
        Lmo[k0:k0 + 128] = Mk.reshape(-1, n * n)[:, flat]
    sample = torch.from_numpy(np.random.default_rng(4).choice(npr, 256, replace=False)).to(dev)
    ref = Lmo[:, sample].T.contiguous() @ Lmo
    got = torch.stack([out64[int(q) * npr:(int(q) + 1) * npr] for q in sample.tolist()])
    cf_err = float((ref - got).abs().max())
    out = {
        "metric": "AO->MO four-index transform time (FP64 MFMA vs BF16x3 MFMA)",
        "value": res["fp64"]["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"ao2mo_{n}: n_ao = n_mo = {n}, {nvec} Cholesky vectors N(0, 0.05^2), "
                               f"8-fold packed AO input, 4-fold packed MO output", "npair": npr,
                   "alg_flops": f_alg},
        "fp64": res["fp64"], "bf16x3": res["bf16x3"],
        "bf16x3_max_abs_err_vs_fp64": dmax, "max_abs_mo_integral": amax,
        "fp64_max_abs_err_vs_closed_form_256_columns": cf_err,
        "roofline": {"bound": "mfma", "achieved": res["fp64"]["tflops"], "peak": 78.6, "unit": "TFLOP/s",
                     "frac": res["fp64"]["frac_of_peak"]},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def cpu_baseline_fci16(prob, sigma_vectors, threads):
    """Oracle sigma of 4 vectors on the determinants i = 0 mod 256 of the 16-orbital 8a8b space
    (647,020 rows), less the fixed cost of a call (its per-thread private sigma of N x 4 values,
    zeroed and then summed, as the reference's singleshot sigma does), measured by a call on one
    row; the per-row cost, scaled to all rows, is one full sigma; times the sigma vectors of the
    GPU FP64 solve / 4 gives one CPU solve."""
    import numpy as np

    from oracle import refcpu

    k = 4
    t0 = time.perf_counter()
    oc = refcpu.OracleCI(prob)
    t_set = time.perf_counter() - t0
    C = np.random.default_rng(0).standard_normal((oc.N, k))
    stride = 256
    t0 = time.perf_counter()
    oc.sigma_sample(C, oc.N)  # row 0 only: the fixed cost
    t_fixed = time.perf_counter() - t0
    t0 = time.perf_counter()
    oc.sigma_sample(C, stride)
    t_sample = time.perf_counter() - t0
    rows = -(-oc.N // stride)
    if t_sample <= 1.5 * t_fixed:
        raise RuntimeError(f"sample ({t_sample:.1f} s) not above the fixed cost ({t_fixed:.1f} s)")
    per_row = (t_sample - t_fixed) / (rows - 1)
    t_sigma = per_row * oc.N
    t_solve = t_sigma * sigma_vectors / k
    conn = 5792
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"extrapolated: oracle symmetric-scatter sigma (k={k} vectors) on determinants i = 0 mod "
                   f"{stride} ({rows} of {oc.N}): {t_sample:.1f} s, less {t_fixed:.1f} s of fixed per-call cost "
                   f"(one row) = {per_row * 1e6:.2f} us per row on {threads} threads = "
                   f"{conn * k / per_row / 1e9:.3f} G nonzero-vector products/s; full sigma {t_sigma:.0f} s x "
                   f"{sigma_vectors / k:.1f} sigma applications of the GPU FP64 solve ({sigma_vectors} vectors); "
                   f"oracle determinant set {t_set:.1f} s, not counted"),
        "sigma_s": t_sigma,
        "per_nonzero_ns": per_row / (conn * k) * 1e9,
    }


def bench_fci16(args):
    """Electronic FCI 16 orbitals 8a8b (165,636,900 determinants), 4 roots, residual <= 1e-6:
    FP64 Davidson an iteration at a time (W warm-up + K timed iterations, then to convergence),
    the mixed-precision solve beside it, the link-GEMM roofline and the oracle CPU baseline."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet

    assert torch.cuda.is_available(), "no GPU visible"
    torch.zeros(1, device="cuda")
    prob = wl.config3()
    ctx = Context(0)
    t0 = time.perf_counter()
    ds = DetSet(prob, ctx)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    N, nev = ds.N_dets, prob.nroots
    log(f"[fci16] N={N} roots={nev} setup {setup:.2f}s")
    opts = dict(nvec_init=args.fci_nvec, nvec_max=args.fci_nmax, tol=1e-6, maxit=args.fci_maxit,
                want_vectors=False, perturb=args.fci_perturb)

    def solve(prec, warm=0, timed=0):
        """Whole solve; with timed > 0, iterations warm+1 .. warm+timed are timed (each
        iteration ends with the solve's host synchronisation)."""
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        sv = ds.davidson_solver(nev, precision=prec, **opts)
        it, t_a, t_b = 0, None, None
        done = False
        while not done:
            if timed and it == warm:
                torch.cuda.synchronize()
                t_a = time.perf_counter()
            done = sv.step()
            it += 1
            if timed and it == warm + timed:
                torch.cuda.synchronize()
                t_b = time.perf_counter()
            if it % 25 == 0:
                log(f"[fci16] {'fp64' if prec == PQG_FP64 else 'mixed'} iteration {it} "
                    f"({time.perf_counter() - t_start:.1f}s)")
        E, _, rep = sv.result()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t_start
        sv.close()
        steps_ms = None
        if timed:
            if t_b is None:
                raise RuntimeError(f"the solve finished after {it} iterations, before {warm} + {timed}")
            steps_ms = (t_b - t_a) * 1e3 / timed
        res = {"seconds": round(dt, 3), "iterations": rep.iterations, "sigma_vectors": rep.sigma_calls,
               "sigma_vectors_fp64": rep.sigma_calls_fp64, "energies": [float(x) for x in E],
               "residuals": [float(rep.residual[i]) for i in range(nev)]}
        return res, steps_ms, E

    fp64, ms_per_step, E64 = solve(PQG_FP64, args.warmup, args.steps)
    log(f"[fci16] fp64: {fp64['seconds']:.1f}s {fp64['iterations']} iterations, {ms_per_step:.1f} ms per timed "
        f"iteration, E={E64}")
    mixed, _, Em = solve(PQG_MIXED)
    dE = [float(abs(a - b)) for a, b in zip(Em, E64)]
    mixed["speedup_vs_fp64"] = round(fp64["seconds"] / mixed["seconds"], 3)
    mixed["abs_dE_vs_fp64_per_root"] = dE
    log(f"[fci16] mixed: {mixed['seconds']:.1f}s {mixed['iterations']} iterations, |dE| {max(dE):.2e}")

    # roofline: one FP64 and one FP32 sigma of 4 device-resident vectors with kernel events
    k = 4
    sp = prob.species[0]
    n, na = sp.n_orb, sp.n_alpha
    n_s = math.comb(n, na)
    s1 = na * (n - na)
    d2 = math.comb(na, 2) * math.comb(n - na, 2)
    conn = 2 * s1 + 2 * d2 + s1 * s1
    f_alg = 2.0 * k * N * (conn + 1)
    gemm_flops = k * (2.0 * n_s * s1 * (n * (n - 1) // 2 + n + 1 + s1) * n_s + 2.0 * n_s * s1 * (n + 1 + s1) * n_s)
    C = torch.randn(N * k, dtype=torch.float64, device="cuda") / math.sqrt(N)
    S = torch.empty_like(C)
    kern = {}
    # the mixed-precision sigma runs its link GEMM as BF16x3 (three bf16 products per FP32 product on
    # the bf16 matrix cores: 2.5 PFLOP/s dense / 3 of FP32-equivalent work)
    # the gather (k_lg_gather, HBM-bound) reads U once per pass (N x links x vectors elements) and
    # writes sigma in both passes, reading it back in the accumulating one
    links = s1
    for tag, prec, suf, peak in (("fp64", PQG_FP64, "_f64", 78.6), ("bf16x3", PQG_MIXED, "_f32", 2500.0 / 3)):
        w = 8 if prec == PQG_FP64 else 4
        gather_bytes = k * N * w * (2 * links + 3)
        ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, prec)  # warm-up
        torch.cuda.synchronize()
        ctx.profile(True)
        for nm in ("lg_gemm", "lg_gather"):
            ctx.profile_read(nm + suf)
        t0 = time.perf_counter()
        ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ctx.profile(False)
        launches, gemm_ms, _ = ctx.profile_read("lg_gemm" + suf)
        _, gather_ms, _ = ctx.profile_read("lg_gather" + suf)
        kern[tag] = {"sigma_ms_4_vectors": round(dt * 1e3, 2), "k_lg_u_ms": round(gemm_ms, 2), "launches": launches,
                     "avg_launch_ms": round(gemm_ms / max(launches, 1), 4), "gather_ms": round(gather_ms, 2),
                     "alg_tflops": round(f_alg / (gemm_ms * 1e-3) / 1e12, 3),
                     "alg_frac": round(f_alg / (gemm_ms * 1e-3) / 1e12 / peak, 4),
                     "gemm_tflops": round(gemm_flops / (gemm_ms * 1e-3) / 1e12, 3),
                     "gemm_frac": round(gemm_flops / (gemm_ms * 1e-3) / 1e12 / peak, 4), "peak_tflops": peak,
                     "gather_alg_bytes": gather_bytes,
                     "gather_hbm_tbs": round(gather_bytes / (gather_ms * 1e-3) / 1e12, 3),
                     "gather_hbm_frac": round(gather_bytes / (gather_ms * 1e-3) / 8e12, 4)}
        log(f"[fci16] sigma {tag}: {kern[tag]}")
    del C, S
    r64 = kern["fp64"]
    roofline = {"bound": "mfma", "achieved": r64["alg_tflops"], "peak": 78.6, "unit": "TFLOP/s",
                "frac": r64["alg_frac"], "traffic": traffic_from_profiles("lg_u_f64"),
                "kernel": "k_lg_u<double> (link GEMM, FP64 MFMA)", "avg_launch_ms": r64["avg_launch_ms"],
                "launches_per_sigma": r64["launches"],
                "alg_flops_per_launch": int(f_alg / max(r64["launches"], 1)),
                "gemm_flops_per_launch": int(gemm_flops / max(r64["launches"], 1)),
                "gemm_tflops": r64["gemm_tflops"], "gemm_frac": r64["gemm_frac"]}

    cpu = None
    if not args.no_cpu:
        try:
            ds.close()
            ds = None
            torch.cuda.empty_cache()
            cpu = cpu_baseline_fci16(prob, fp64["sigma_vectors"], args.cpu_threads)
            log(f"[fci16] cpu baseline: {cpu}")
            cpu.pop("sigma_s", None)
        except Exception as e:  # the CPU leg is a reported baseline, not the product
            log(f"[fci16] cpu baseline skipped: {e}")
    out = {
        "metric": "Davidson time-to-1e-6 Eh (FP64 vs mixed); CI solves/s + p99 latency",
        "value": round(1.0 / fp64["seconds"], 6),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"fci_16o_8a8b: electronic full CI, 16 orbitals, 8a8b electrons, {N} determinants, "
                                f"{nev} roots, Davidson residual <= 1e-6; step = one FP64 Davidson iteration"),
                   "n_dets": N, "roots": nev, "setup_s": round(setup, 3), "num_subspace_vec": args.fci_nvec,
                   "max_subspace": args.fci_nmax, "start_perturbation": args.fci_perturb,
                   "fp64_time_to_1e-6_s": fp64["seconds"], "mixed_time_to_1e-6_s": mixed["seconds"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64": fp64,
        "mixed": mixed,
        "sigma_kernels": kern,
    }
    print(json.dumps(out), flush=True)
    if ds is not None:
        ds.close()
    ctx.close()


def _serve_sequential(args, reqs, ctx, nvec):
    """Sequential plain FP64 solves, one request at a time: per-request seconds and energies."""
    from polyquant_amd._lib import PQG_FP64
    from polyquant_amd.ci import DetSet

    n = len(reqs)
    seq_t, seq_E, seq_fail = [], [], 0
    t_all = time.perf_counter()
    for i, (_, prob) in enumerate(reqs):
        t0 = time.perf_counter()
        try:
            ds = DetSet(prob, ctx)
            E, _, _ = ds.davidson(prob.nroots, nvec_init=nvec, tol=1e-6, precision=PQG_FP64, want_vectors=False,
                                  perturb=args.serve_perturb)
            ds.close()
            seq_E.append([float(x) for x in E])
        except Exception as e:  # a failed request is counted, the loop goes on
            seq_E.append(None)
            seq_fail += 1
            log(f"[serve] sequential request {i} failed: {e}")
        seq_t.append(time.perf_counter() - t0)
        log(f"[serve] sequential {i + 1}/{n}: {prob.n_dets_fci()} determinants, {prob.nroots} roots, "
            f"{seq_t[-1]:.3f}s (total {time.perf_counter() - t_all:.1f}s)")
    return {"total_s": time.perf_counter() - t_all, "seconds": seq_t, "energies": seq_E, "failed": seq_fail}


def _serve_loop(args, reqs, ctx, nvec):
    """Replays the arrival times into the serving loop (mixed precision, FP64 correction)."""
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.serve import Server

    n = len(reqs)
    srv = Server(ctx, max_slots=args.serve_slots)
    got, ids = {}, {}
    t_start = time.perf_counter()
    nxt = 0
    while len(got) < n:
        now = time.perf_counter() - t_start
        while nxt < n and reqs[nxt][0] <= now:
            rid = srv.submit(reqs[nxt][1], nvec_init=nvec, tol=1e-6, precision=PQG_MIXED, perturb=args.serve_perturb)
            ids[rid] = nxt
            nxt += 1
        for r in srv.poll():
            got[ids[r["id"]]] = r
            log(f"[serve] served {len(got)}/{n}: request {ids[r['id']]} status {r['status']}, "
                f"latency {r['t_done'] - r['t_submit']:.3f}s ({time.perf_counter() - t_start:.1f}s)"
                + (f" {r['message']}" if r["status"] else ""))
        time.sleep(0.0005)
    makespan = time.perf_counter() - t_start
    srv.close()
    return {"makespan_s": makespan,
            "latency_s": [got[i]["t_done"] - got[i]["t_submit"] for i in range(n)],
            "status": [int(got[i]["status"]) for i in range(n)],
            "energies": [[float(x) for x in got[i]["energies"]] for i in range(n)]}


def _merge_legs(parts):
    """Sequential legs run over consecutive request ranges, concatenated."""
    out = {"total_s": 0.0, "seconds": [], "energies": [], "failed": 0}
    for p in parts:
        out["total_s"] += p["total_s"]
        out["seconds"] += p["seconds"]
        out["energies"] += p["energies"]
        out["failed"] += p["failed"]
    return out


def _serve_report(args, n, seq, srv):
    import numpy as np

    lat = np.array(srv["latency_s"])
    solved = sum(1 for i in range(n) if srv["status"][i] == 0)
    ok = [i for i in range(n) if srv["status"][i] == 0 and seq["energies"][i] is not None]
    dE = max((float(np.max(np.abs(np.array(srv["energies"][i]) - np.array(seq["energies"][i])))) for i in ok),
             default=0.0)
    seq_lat = np.array(seq["seconds"])
    makespan = srv["makespan_s"]
    return {
        "metric": "CI solves/s + p99 latency (serving loop vs sequential plain FP64)",
        "value": round(solved / makespan, 3), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "mixed (FP32 sigma, FP64 correction)", "data": "synthetic",
        "config": {"workload": f"serving: {n} CI requests, Poisson arrivals 200/s, n_orb 6-14, alpha/beta 2..n/2, "
                               "positron 0/1 with 6-12 orbitals, 1-4 roots, HBM budget 200 GB",
                   "max_slots": args.serve_slots, "num_subspace_vec": args.serve_nvec,
                   "start_perturbation": args.serve_perturb, "admission": "smallest work first, oldest first after 30 s"},
        "serving": {"makespan_s": round(makespan, 3), "solves_per_s": round(solved / makespan, 3),
                    "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                    "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                    "failed": int(sum(1 for i in range(n) if srv["status"][i] != 0))},
        "sequential_fp64": {"total_s": round(seq["total_s"], 3), "solves_per_s": round(n / seq["total_s"], 3),
                            "p50_solve_s": round(float(np.percentile(seq_lat, 50)), 4),
                            "p99_solve_s": round(float(np.percentile(seq_lat, 99)), 4), "failed": seq["failed"]},
        "speedup_solves_per_s": round((solved / makespan) / ((n - seq["failed"]) / seq["total_s"]), 3),
        "max_abs_dE_serving_vs_sequential": dE,
    }


def bench_serve(args):
    """Serving loop: 512 CI requests with Poisson arrivals (200/s), mixed precision, vs sequential FP64.

    --serve-leg sequential|serving runs one leg and writes it to --serve-save (JSON); --serve-combine
    SEQ SRV prints the report of two saved legs.  The default runs both legs in one process."""
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd.ci import Context

    reqs = wl.serving_requests()[: args.serve_n]
    n = len(reqs)
    # num_subspace_vec of every request (both legs); the largest subspace is the reference's
    # max(2 * num_subspace_vec, 10 * roots) (epci.cpp:589-608)
    nvec = args.serve_nvec
    if args.serve_combine:
        seq = _merge_legs([json.load(open(f)) for f in args.serve_combine[0].split(",")])
        srv = json.load(open(args.serve_combine[1]))
        print(json.dumps(_serve_report(args, n, seq, srv)), flush=True)
        return
    assert torch.cuda.is_available(), "no GPU visible"
    ctx = Context(0, hbm_budget=200 << 30)
    seq = srv = None
    if args.serve_leg in ("both", "sequential"):
        lo, hi = args.serve_range if args.serve_range else (0, n)
        seq = _serve_sequential(args, reqs[lo:hi], ctx, nvec)
    if args.serve_leg in ("both", "serving"):
        srv = _serve_loop(args, reqs, ctx, nvec)
    ctx.close()
    if args.serve_save:
        json.dump(seq if args.serve_leg == "sequential" else srv, open(args.serve_save, "w"))
    if seq is not None and srv is not None:
        print(json.dumps(_serve_report(args, n, seq, srv)), flush=True)
    else:
        leg = seq if seq is not None else srv
        total = leg.get("total_s", leg.get("makespan_s"))
        print(json.dumps({"serve_leg": args.serve_leg, "requests": n, "seconds": round(total, 3)}), flush=True)


def bench_sigma(args):
    """Sigma throughput of the link-GEMM path: 16 orbitals 8a8b (165,636,900 determinants) or
    water/6-31G (1,656,369), k = 4 vectors per application, FP64 and mixed (BF16x3), with the roofline of
    the link-GEMM kernel.  Algorithmic flops of one k_lg_u launch = 2 * rows * L * K * columns
    (M = L links per row string, K = the column set's off-diagonal and diagonal pairs + the own
    row + L, the GEMM as posed, without padding); launch durations from HIP events on the CI
    stream."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet

    assert torch.cuda.is_available(), "no GPU visible"
    prob = wl.config3() if args.sigma_workload == "fci16" else wl.water_problem("6-31g")
    ctx = Context(0)
    ds = DetSet(prob, ctx)
    N, k = ds.N_dets, 4
    sp = prob.species[0]
    n, na, nb = sp.n_orb, sp.n_alpha, sp.n_beta
    n_a, n_b = math.comb(n, na), math.comb(n, nb)
    La, Lb = na * (n - na), nb * (n - nb)
    # GEMM flops of one application per vector: pass A (rows alpha; K = beta off-diagonal and
    # diagonal pairs, the own row, the alpha links) and pass B (rows beta; K = alpha diagonal
    # pairs, the own row, the beta links)
    gemm_flops = (2.0 * n_a * La * (n * (n - 1) // 2 + n + 1 + La) * n_b
                  + 2.0 * n_b * Lb * (n + 1 + Lb) * n_a)
    C = torch.randn(N * k, dtype=torch.float64, device="cuda") / math.sqrt(N)
    S = torch.empty_like(C)
    torch.cuda.synchronize()
    names = ["lg_gemm", "lg_gather"]
    res = {}
    # the mixed-precision sigma runs its link GEMM as BF16x3 (three bf16 products per FP32
    # product: 2.5 PFLOP/s dense bf16 / 3 of FP32-equivalent work)
    for tag, prec in (("fp64", PQG_FP64), ("bf16x3", PQG_MIXED)):
        suf = "_f64" if prec == PQG_FP64 else "_f32"
        for _ in range(args.warmup):
            ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, prec)
        ctx.profile(True)
        for nm in names:
            ctx.profile_read(nm + suf)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, prec)
        dt = (time.perf_counter() - t0) / args.steps
        ctx.profile(False)
        prof = {nm: ctx.profile_read(nm + suf) for nm in names}
        launches, gemm_ms, _ = prof["lg_gemm"]
        gemm_s = gemm_ms * 1e-3 / args.steps
        peak = 78.6 if prec == PQG_FP64 else 2500.0 / 3
        achieved = gemm_flops * k / gemm_s / 1e12
        res[tag] = {"ms_per_sigma_4_vectors": round(dt * 1e3, 3),
                    "det_vectors_per_s": round(N * k / dt, 1),
                    "lg_gemm_ms": round(gemm_s * 1e3, 3), "lg_gemm_launches_per_sigma": launches // args.steps,
                    "lg_gemm_tflops": round(achieved, 3), "lg_gemm_frac_of_peak": round(achieved / peak, 4),
                    "peak_tflops": peak,
                    "lg_gather_ms": round(prof["lg_gather"][1] / args.steps, 3)}
        log(f"[sigma] {tag}: {res[tag]}")
    fp64 = res["fp64"]
    out = {
        "metric": "sigma (H C) throughput, link-GEMM path", "value": fp64["det_vectors_per_s"],
        "unit": "determinant-vectors/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic" if args.sigma_workload == "fci16" else "water 6-31G",
        "config": {"workload": (f"{args.sigma_workload}: {n} orbitals, {na}a{nb}b, {N} determinants, "
                                f"k = {k} vectors per sigma"), "gemm_flops_per_vector": gemm_flops},
        "fp64": fp64, "bf16x3": res["bf16x3"],
        "bf16x3_speedup": round(fp64["ms_per_sigma_4_vectors"] / res["bf16x3"]["ms_per_sigma_4_vectors"], 3),
        "roofline": {"bound": "mfma", "achieved": fp64["lg_gemm_tflops"], "peak": 78.6, "unit": "TFLOP/s",
                     "frac": fp64["lg_gemm_frac_of_peak"], "traffic": None, "kernel": "k_lg_u<double>"},
    }
    print(json.dumps(out), flush=True)
    ds.close()
    ctx.close()


def bench_rdm(args):
    """One-particle density matrices (create_1rdm) of k = 4 vectors of the 16-orbital 8a8b space
    (165,636,900 determinants), alpha and beta string sets.  k_rdm_group is HBM-bound: its
    algorithmic bytes per launch are the vectors read once (N * 8 * k) plus the group Gram
    entries it writes (C(16, 7) groups x 45 x k x 8); launch durations from HIP events on the CI
    stream."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd.ci import Context, DetSet

    assert torch.cuda.is_available(), "no GPU visible"
    prob = wl.config3()
    ctx = Context(0)
    ds = DetSet(prob, ctx)
    sp = prob.species[0]
    N, k, n = ds.N_dets, 4, sp.n_orb
    n_t, L = math.comb(n, sp.n_alpha), sp.n_alpha * (n - sp.n_alpha)
    C = torch.randn(N * k, dtype=torch.float64, device="cuda") / math.sqrt(N)
    D = torch.empty(k * n * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def step():
        for spin in (0, 1):
            ds.create_1rdm_device(C.data_ptr(), N, k, 0, spin, D.data_ptr())

    for _ in range(args.warmup):
        step()
    ctx.profile(True)
    for nm in ("rdm_group", "rdm_links"):
        ctx.profile_read(nm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.profile(False)
    launches, kern_ms, _ = ctx.profile_read("rdm_group")
    kernel = "k_rdm_group<9>"
    m = n - sp.n_alpha + 1
    out_bytes = math.comb(n, sp.n_alpha - 1) * (m * (m + 1) // 2) * k * 8
    if launches == 0:
        launches, kern_ms, _ = ctx.profile_read("rdm_links")
        kernel = "k_rdm_links"
        out_bytes = n_t * (L + 1) * k * 8
    per_launch_s = kern_ms * 1e-3 / launches
    alg_bytes = N * 8 * k + out_bytes
    achieved = alg_bytes / per_launch_s / 1e9
    log(f"[rdm] {dt * 1e3:.2f} ms per step (2 string sets x {k} vectors), {kernel} "
        f"{per_launch_s * 1e3:.3f} ms/launch, {achieved:.0f} GB/s")
    # <S^2> of the same vectors (evaluate_s2): boundary permutation + k_s2
    ctx.profile(True)
    ctx.profile_read("s2")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s2v = ds.evaluate_s2_device(C.data_ptr(), N, k)
    s2_dt = (time.perf_counter() - t0) / args.steps
    ctx.profile(False)
    s2_launches, s2_ms, _ = ctx.profile_read("s2")
    s2_launch_s = s2_ms * 1e-3 / s2_launches
    s2 = {"ms_per_call_4_vectors": round(s2_dt * 1e3, 3), "kernel": "k_s2p",
          "kernel_ms": round(s2_launch_s * 1e3, 3),
          "kernel_algorithmic_GBps": round(N * 8 * k / s2_launch_s / 1e9, 1),
          "values": [round(float(x), 6) for x in s2v[:, 0]]}
    log(f"[rdm] S^2: {s2}")
    cpu = None
    if not args.no_cpu:
        from oracle import refcpu

        t0 = time.perf_counter()
        oc = refcpu.OracleCI(prob)
        log(f"[rdm] oracle determinant set {time.perf_counter() - t0:.1f}s")
        c0 = C[:N].cpu().numpy()
        s_big, s_small = 4096, 512
        t0 = time.perf_counter()
        oc.rdm1_sample(c0, 0, 0, s_big)
        t_big = time.perf_counter() - t0
        t0 = time.perf_counter()
        oc.rdm1_sample(c0, 0, 0, s_small)
        t_small = time.perf_counter() - t0
        rows_big, rows_small = -(-N // s_big), -(-N // s_small)
        per_det = max((t_small - t_big) / (rows_small - rows_big), 1e-12)
        t_step = per_det * N * 2 * k
        cpu = {"value": round(N * 2 * k / t_step, 1), "unit": "determinant-vectors/s", "cores": args.cpu_threads,
               "kind": "port",
               "sample": (f"oracle create_1rdm on determinants i = 0 mod {s_small} ({rows_small} of {N}) and mod "
                          f"{s_big}, one vector, alpha set; {per_det * 1e6:.3f} us per determinant on "
                          f"{args.cpu_threads} threads -> {t_step:.1f} s for 2 sets x {k} vectors")}
        log(f"[rdm] cpu {cpu}")
    out = {
        "metric": "one-particle density matrix throughput", "value": round(N * 2 * k / dt, 1),
        "unit": "determinant-vectors/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"fci_16o_8a8b: 1-RDM of the alpha and beta string sets, {N} determinants, "
                                f"k = {k} vectors per call")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4),
                     "traffic": traffic_from_profiles("rdm_group") if kernel.startswith("k_rdm_group") else None,
                     "kernel": kernel,
                     "algorithmic_bytes_per_launch": alg_bytes, "ms_per_launch": round(per_launch_s * 1e3, 4)},
        "cpu_baseline": cpu, "s2": s2,
    }
    print(json.dumps(out), flush=True)
    ds.close()
    ctx.close()


def bench_h2o(args):
    """Water/6-31G full CI (13 orbitals, 5a5b, 1,656,369 determinants), RHF orbitals from the CPU."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet

    assert torch.cuda.is_available(), "no GPU visible"
    t0 = time.perf_counter()
    prob = wl.water_problem("6-31g", nroots=1)
    t_scf = time.perf_counter() - t0
    ctx = Context(0)
    ds = DetSet(prob, ctx)
    N = ds.N_dets
    log(f"[h2o] RHF {prob.e_rhf:.10f} ({t_scf:.1f}s on the CPU), N={N}")
    res = {}
    for name, prec in (("fp64", PQG_FP64), ("mixed", PQG_MIXED)):
        ds.davidson(1, nvec_init=args.h2o_nvec, tol=1e-6, precision=prec, want_vectors=False)  # warm-up
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            E, _, rep = ds.davidson(1, nvec_init=args.h2o_nvec, tol=1e-6, precision=prec, want_vectors=False)
            times.append(time.perf_counter() - t0)
        res[name] = {"seconds": round(min(times), 4), "iterations": rep.iterations, "sigma_vectors": rep.sigma_calls,
                     "energy": float(E[0] + prob.e_shift), "residual": float(rep.residual[0])}
        log(f"[h2o] {name}: {min(times) * 1e3:.1f} ms, {rep.iterations} iterations, E={E[0] + prob.e_shift:.10f}")
    # one FP64 sigma with kernel events
    ctx.profile(True)
    for name in ("sigma_f64",):
        ctx.profile_read(name)
    C = np.random.default_rng(0).standard_normal((N, 1))
    ds.sigma(C)
    ctx.profile(False)
    launches, sig_ms, _ = ctx.profile_read("sigma_f64")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, res["fp64"]["sigma_vectors"], 1, args.cpu_threads)
        cpu.pop("sigma_s", None)
    out = {
        "metric": "Davidson time-to-1e-6 Eh (FP64 vs mixed); CI solves/s",
        "value": round(1.0 / res["fp64"]["seconds"], 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": True, "dtype": "f64", "data": "water 6-31G, RHF orbitals from the CPU integral code",
        "config": {"workload": f"h2o_631g_fci: H2O (O-H 0.9572 A, HOH 104.52 deg), 6-31G, 13 orbitals, 5a5b, "
                               f"{N} determinants, 1 root, residual <= 1e-6", "e_rhf": prob.e_rhf,
                   "num_subspace_vec": args.h2o_nvec},
        "fp64": res["fp64"], "mixed": res["mixed"],
        "mixed_speedup": round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 3),
        "abs_dE_mixed_vs_fp64": abs(res["mixed"]["energy"] - res["fp64"]["energy"]),
        "sigma_ms_fp64_one_vector": round(sig_ms / max(launches, 1), 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ds.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--workload", default="fci16", choices=["fci16", "ep48", "ao2mo", "serve", "h2o", "sigma", "rdm"],
                    help="fci16 (default): 16o 8a8b FP64 vs mixed Davidson; ep48: ep-FCI 48/48/48 solves; "
                         "ao2mo: n = 320 transform")
    ap.add_argument("--h2o-nvec", type=int, default=5)
    ap.add_argument("--sigma-workload", default="fci16", choices=["fci16", "h2o"])
    ap.add_argument("--serve-n", type=int, default=512)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-leg", default="both", choices=["both", "sequential", "serving"])
    ap.add_argument("--serve-range", nargs=2, type=int, default=None, metavar=("LO", "HI"),
                    help="sequential leg over requests [LO, HI) only")
    ap.add_argument("--serve-nvec", type=int, default=8)
    ap.add_argument("--serve-perturb", type=float, default=1e-2,
                    help="start-vector perturbation of every request (both legs); 0 = unit vectors")
    ap.add_argument("--serve-save", default=None, help="write the leg's per-request results (JSON)")
    ap.add_argument("--serve-combine", nargs=2, default=None, metavar=("SEQ", "SRV"),
                    help="print the serving report from saved legs (SEQ: comma-separated sequential parts)")
    ap.add_argument("--dav-nvec", type=int, default=16, help="num_subspace_vec of the 48+48 workload")
    ap.add_argument("--dav-nmax", type=int, default=40, help="largest search space of the 48+48 workload")
    ap.add_argument("--fci-nvec", type=int, default=16)
    ap.add_argument("--fci-nmax", type=int, default=48)
    ap.add_argument("--fci-maxit", type=int, default=400)
    ap.add_argument("--fci-perturb", type=float, default=1e-2,
                    help="start-vector perturbation of both solves; 0 = unit vectors")
    ap.add_argument("--ao-n", type=int, default=320)
    ap.add_argument("--ao-reps", type=int, default=2)
    ap.add_argument("--ao-budget-gb", type=float, default=120.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer (the known-size kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the CI path is built and "
              f"measured on a single MI355X.", file=sys.stderr)
        sys.exit(2)
    # the oracle's OpenMP runtime reads this when it is first loaded
    os.environ["OMP_NUM_THREADS"] = str(args.cpu_threads)
    if args.copy_calibration:
        import torch

        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 8, "launches": 3}))
        return
    if args.workload == "ao2mo":
        bench_ao2mo(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "h2o":
        bench_h2o(args)
        return
    if args.workload == "rdm":
        bench_rdm(args)
        return
    if args.workload == "sigma":
        bench_sigma(args)
        return
    if args.workload == "ep48":
        bench_ep48(args)
        return
    bench_fci16(args)


def bench_ep48(args):
    """Electron + positron full CI 48 + 48 orbitals (110,592 determinants), 3 roots: whole FP64
    solves as steps, the pair-GEMM roofline, the mixed solve beside it."""
    import numpy as np
    import torch

    from polyquant_amd import workloads as wl
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet

    assert torch.cuda.is_available(), "no GPU visible"
    torch.zeros(1, device="cuda")
    prob = wl.config2()
    ctx = Context(0)
    t0 = time.perf_counter()
    ds = DetSet(prob, ctx)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    N = ds.N_dets
    nev = prob.nroots
    log(f"[bench] workload ep_fci_48e_48p: N={N} roots={nev} setup {setup_s:.3f}s")

    # num_subspace_vec = 16 and a 40-vector search space (the reference's knobs,
    # epci.cpp:589-608; its defaults 5 / max(10, 10*nstates) need ~2.5x more iterations here).
    # Measured on one MI355X, whole FP64 solves: 84.4 ms at 16/40, 92.0 ms at 16/64, 95.4 ms at
    # 16/32 (the reference's max(2*16, 10*3)), 91.2 ms at 8/32 -- fewer iterations with a larger
    # space stop paying once the subspace eigensolve grows with m^3.
    nvec_init, nvec_max = args.dav_nvec, args.dav_nmax

    def solve(precision):
        return ds.davidson(nev, nvec_init=nvec_init, nvec_max=nvec_max, tol=1e-6, precision=precision,
                           want_vectors=False)

    for _ in range(args.warmup):
        solve(PQG_FP64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = []
    for _ in range(args.steps):
        E, _, rep = solve(PQG_FP64)
        reps.append(rep)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = args.steps / dt
    iters = reps[-1].iterations

    # one more solve with per-launch HIP events (outside the timed region)
    kernels = ("tensor_gemm_f64", "sigma_pass_nc3_f64", "sigma_pass_nc1_f64", "sigma_f64")
    ctx.profile(True)
    for name in kernels:
        ctx.profile_read(name)
    solve(PQG_FP64)
    torch.cuda.synchronize()
    ctx.profile(False)
    prof = {name: ctx.profile_read(name) for name in kernels}
    s_launch, s_ms, s_vectors = prof["sigma_f64"]
    sigma_calls = s_launch
    log(f"[bench] fp64: {ms_per_step:.1f} ms/solve, {iters} iterations, {sigma_calls} sigma calls, E={E}")
    log(f"[bench] kernel profile: {prof}")

    # roofline of the dominant sigma kernel
    dominant = max(kernels[:3], key=lambda n: prof[n][1])
    launches, kern_ms, units = prof[dominant]
    avg_s = kern_ms * 1e-3 / max(launches, 1)
    traffic = traffic_from_profiles(dominant)
    if dominant == "tensor_gemm_f64":
        # algorithmic flops of the sigma calls (one pair-GEMM launch per sigma call)
        conn = 3 * 47 + 3 * 47 * 47
        alg_flops = 2.0 * N * (conn + 1) * s_vectors
        achieved = alg_flops / max(kern_ms * 1e-3, 1e-30) / 1e12
        peak = 78.6
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(achieved / peak, 5), "traffic": traffic,
                    "kernel": "k_pair_gemm<double,9>", "avg_launch_ms": round(avg_s * 1e3, 5),
                    "alg_flops_per_launch": int(alg_flops / max(launches, 1)), "launches": launches}
    else:
        npair = 48 * 49 // 2
        coupling = 3 * npair * npair * 8
        alg_bytes = 2 * 8 * N * units + 8 * N * launches + coupling * launches
        achieved = alg_bytes / max(launches, 1) / avg_s / 1e9
        peak = 8000.0
        roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": peak, "unit": "GB/s",
                    "frac": round(achieved / peak, 6), "traffic": traffic,
                    "kernel": "k_sigma_pass<double,3>", "avg_launch_ms": round(avg_s * 1e3, 4),
                    "alg_bytes_per_launch": int(alg_bytes / max(launches, 1)), "launches": launches}

    # mixed precision beside it
    for _ in range(args.warmup):
        solve(PQG_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Em, _, rep_m = solve(PQG_MIXED)
    torch.cuda.synchronize()
    dtm = time.perf_counter() - t0
    mixed = {"ms_per_solve": round(dtm * 1e3 / args.steps, 3), "solves_per_s": round(args.steps / dtm, 4),
             "speedup_vs_fp64": round(dt / dtm, 3), "max_abs_dE_vs_fp64": float(np.max(np.abs(Em - E))),
             "iterations": rep_m.iterations, "sigma_vectors_fp32": rep_m.sigma_calls - rep_m.sigma_calls_fp64,
             "sigma_vectors_fp64": rep_m.sigma_calls_fp64}
    log(f"[bench] mixed: {mixed}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(prob, sigma_calls, nev, args.cpu_threads)
            log(f"[bench] cpu baseline: {cpu}")
            cpu.pop("sigma_s", None)
        except Exception as e:  # the CPU leg is a reported baseline, not the product
            log(f"[bench] cpu baseline skipped: {e}")

    out = {
        "metric": "Davidson time-to-1e-6 Eh (FP64 vs mixed); CI solves/s + p99 latency",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ep_fci_48e_48p: electron+positron FCI, 48+48 orbitals, 1a1b electrons, "
                               "1 positron, 110592 determinants, 3 roots, residual <= 1e-6",
                   "n_dets": N, "roots": nev, "setup_s": round(setup_s, 4),
                   "num_subspace_vec": nvec_init, "max_subspace": nvec_max,
                   "davidson_iterations": iters, "sigma_calls": sigma_calls,
                   "sigma_ms_avg": round(s_ms / max(s_launch, 1), 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)
    ds.close()
    ctx.close()




if __name__ == "__main__":
    main()
