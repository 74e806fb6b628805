"""Benchmark: sync-free lower-triangular solve on MI355X.

Workload (one "step"): the FP64 forward solve L x = b of the wide-level
matrix -- n = 2,000,000 rows in 200 levels of 10,000, each row 1-2
off-diagonals drawn from the rows of the previous 1-3 levels, off-diagonals
U[-1,1], diagonal 1 + sum|offdiag|, b U[-1,1] -- with the matrix, b and x
resident in HBM.  A step is the per-solve reset plus the solve kernel.

Prints one JSON line:
  value        GFLOPS = 2 * nnzL / step time (nnzL includes the diagonal)
  roofline     solve kernel alone: algorithmic bytes (the reference's
               dataSize) / median MALL-flushed kernel time vs 8 TB/s (warm
               beside it); traffic from the committed rocprofv3 counters
  cpu_baseline the reference's serial executor (serialref.h, compiled from
               its own sources into oracle/_ref) on one host core, same matrix
  extras       FP32 solve, FP32 solve + FP64 refinement to 1e-12 (error vs
               the CPU FP64 result), MALL-flushed (cold) step time

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpTRSV GFLOPS + HBM GB/s; FP32+refinement vs FP64 solve at 1e-12 rel. error"
HBM_PEAK_GBS = 8000.0
WORKLOAD = ("wide-level lower-triangular CSR, n=2,000,000 (200 levels x 10,000 rows, 1-2 off-diagonals "
            "uniform over the previous 1-3 levels), FP64 forward solve, single RHS")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--quick", action="store_true", help="headline workload only (skip the extra workloads)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solve is a single-GPU kernel, measured on one MI355X "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from capellinisptrsv_amd import Plan, api

    dev = torch.device("cuda:0")
    log("generating the wide-level matrix")
    rp, ci, v = api.make_config(1, seed=0)
    n, nnz = len(rp) - 1, len(ci)
    levels = api.matrix_layer(n, n, nnz, rp, ci, v)[0]
    b_h = api.uniform(1, n)
    d_rp, d_ci, d_v = (torch.from_numpy(a).to(dev) for a in (rp, ci, v))
    b = torch.from_numpy(b_h).to(dev)
    x = torch.empty_like(b)

    plan = Plan(d_rp, d_ci, d_v)  # autotuned algorithm / protocol / grid
    info = plan.info()
    log(f"plan: {info}")

    def step():
        plan.solve(b, out=x)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert plan.status() == 0, "a dependency wait gave up"
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    x64 = x.clone()
    # the same solve with the host waiting for each one, as a refined solve
    # does (it reads its convergence test back before returning)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        torch.cuda.synchronize()
    synced_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # the solve kernel alone, HIP events on the stream it is launched on:
    # MALL-warm (back to back, as the reference measures) and MALL-flushed
    # (a 512 MB fill before each launch), medians of >= 20 launches each
    stream = torch.cuda.current_stream()
    flush = torch.empty(512 * 1024 * 1024 // 4, dtype=torch.float32, device=dev)

    def kernel_times(cold):
        ts = []
        for _ in range(max(args.steps, 20)):
            plan.reset(x)
            if cold:
                flush.fill_(1.0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.launch(b, x)
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))
    kernel_ms_warm = kernel_times(False)
    kernel_ms = kernel_times(True)
    assert torch.equal(x, x64)

    # cold step: MALL flushed before every step, each step timed alone
    cold = []
    for _ in range(10):
        flush.fill_(1.0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        e1.synchronize()
        cold.append(e0.elapsed_time(e1))
    del flush

    # spin counter: the same tuning in the counting build (poll loads per row)
    spin = None
    if info["algo"] == "wrtfst" and info["sync"] == "value":
        pc = Plan(d_rp, d_ci, d_v, algo="wrtfst", sync="value", lds_window=info["lds_window"],
                  blocks_per_cu=info["blocks_per_cu"], readiness=info["readiness"], row_threads=info["row_threads"],
                  count_polls=True)
        xc_ = torch.empty_like(b)
        pc.solve(b, out=xc_)
        pc.poll_count()
        polls = []
        for _ in range(5):
            pc.solve(b, out=xc_)
            polls.append(pc.poll_count())
        assert torch.equal(xc_, x64)
        pc.close()
        polls = float(np.median(polls))
        spin = {"polls_per_row": round(polls / n, 3), "polls_per_dependency": round(polls / (nnz - n), 3),
                "polls_per_solve": int(polls)}

    # the reference arithmetic (rounded product, then rounded sum): same tuning
    pr = Plan(d_rp, d_ci, d_v, algo="wrtfst", sync="value", lds_window=info["lds_window"],
              blocks_per_cu=info["blocks_per_cu"], readiness=info["readiness"], row_threads=info["row_threads"],
              arith="reference")
    x_refarith = torch.empty_like(b)
    refarith_ms = _time_gpu(lambda: pr.solve(b, out=x_refarith), args.steps, warmup=3)
    assert pr.status() == 0
    pr.close()

    # FP32 solve, and FP32 solves inside FP64 refinement
    plan32 = Plan(d_rp, d_ci, d_v.float(), algo=info["algo"] if info["algo"] != "csc" else "wrtfst")
    b32, x32 = b.float(), torch.empty(n, dtype=torch.float32, device=dev)
    for _ in range(args.warmup):
        plan32.solve(b32, out=x32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan32.solve(b32, out=x32)
    torch.cuda.synchronize()
    fp32_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    xr = torch.empty_like(b)
    for _ in range(3):
        _, iters, rel = plan.solve_refined(b, tol=1e-12, out=xr)
    torch.cuda.synchronize()
    reps = max(5, args.steps // 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        _, iters, rel = plan.solve_refined(b, tol=1e-12, out=xr)
    torch.cuda.synchronize()
    refined_ms = (time.perf_counter() - t0) * 1e3 / reps

    # algorithmic bytes of one solve (the reference's dataSize,
    # spts_syncfree_cuda_csr_wrt.h:208): row_ptr, col_idx and values read
    # once, a 4-byte readiness word per row, b read once, x written once
    vb = 8
    alg_bytes = 4 * (n + 1) + (4 + vb) * nnz + 4 * n + 2 * vb * n
    ref_datasize = api.datasize(n, n, nnz, vb)  # the reference's formula (includes its int flag array)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    achieved_warm = alg_bytes / (kernel_ms_warm * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_wrtfst_wide_level_2M.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        # only for the kernel configuration those counters were taken on
        if all(t.get("plan", {}).get(key) == info.get(key) for key in ("algo", "sync", "lds_window",
                                                                       "blocks_per_cu", "readiness",
                                                                       "row_threads")) \
                and t.get("n") == n:
            traffic = t.get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(2.0 * nnz / (ms_per_step * 1e6), 3), "unit": "GFLOPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": WORKLOAD, "n": n, "nnzL": nnz, "levels": levels, "algo": info["algo"],
                   "sync": info["sync"], "seed": 0},
        "gbs_algorithmic": round(alg_bytes / (ms_per_step * 1e-3) / 1e9, 2),
        "gbs_reference_formula": round(ref_datasize / (ms_per_step * 1e-3) / 1e9, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_wrtfst" if info["algo"] == "wrtfst" else "k_" + info["algo"],
                     "kernel_ms": round(kernel_ms, 5), "kernel_ms_is": "median, MALL flushed before each launch",
                     "kernel_ms_warm": round(kernel_ms_warm, 5), "achieved_warm": round(achieved_warm, 2),
                     "frac_warm": round(achieved_warm / HBM_PEAK_GBS, 5), "algorithmic_bytes": alg_bytes},
        "per_level_ns": round(kernel_ms * 1e6 / levels, 1),
        "spin_per_row": spin,
        "reference_arith_ms_per_step": round(refarith_ms, 5),
        # latency check: the per-level time 40 % of the HBM roofline would need
        "per_level_ns_for_40pct_roofline": round(alg_bytes / (0.4 * HBM_PEAK_GBS * 1e9) / levels * 1e9, 1),
        "cold_ms_per_step": round(float(np.median(cold)), 5),
        "fp32_ms_per_step": round(fp32_ms, 5),
        "refined_ms_per_solve": round(refined_ms, 5), "refined_iterations": iters,
        "refined_over_fp64": round(refined_ms / ms_per_step, 3),
        "fp64_ms_per_solve_synced": round(synced_ms, 5),
        "refined_over_fp64_synced": round(refined_ms / synced_ms, 3),
    }

    # CPU baseline: the reference's serial executor on the same matrix (1 core)
    xc = None
    if not args.no_cpu:
        import oracle
        cr, cp, cv = oracle.transpose(n, n, nnz, rp, ci, v)
        indeg = oracle.csc_analyser(cr, n, nnz)
        if oracle.ref_available("d64"):
            kind = "reference"

            def cpu_solve():
                return oracle.ref_csc_executor(cp, cr, cv, indeg, n, n, 0, 1, b_h)
        else:
            kind = "port"

            def cpu_solve():
                return oracle.csc_executor(cp, cr, cv, indeg, n, n, 0, 1, b_h)
        xc = cpu_solve()
        times, start = [], time.perf_counter()
        while time.perf_counter() - start < args.cpu_seconds and len(times) < 1000:
            t = time.perf_counter()
            cpu_solve()
            times.append(time.perf_counter() - t)
        cpu_ms = float(np.median(times)) * 1e3
        out["cpu_baseline"] = {"value": round(2.0 * nnz / (cpu_ms * 1e6), 4), "unit": "GFLOPS", "cores": 1,
                               "kind": kind, "ms_per_solve": round(cpu_ms, 3),
                               "sample": f"{len(times)} full solves of the same n=2,000,000 matrix (median), "
                                         f"serial column-driven executor, one core of a {_cpu_model()} "
                                         f"({os.cpu_count()} cores visible)"}
        ref = xc
    else:
        import oracle
        ref = oracle.csr_forward(rp, ci, v, b_h, fused=False)
    den = float(np.max(np.abs(ref)))
    out["fp64_err_vs_cpu"] = float(np.max(np.abs(x64.cpu().numpy() - ref)) / den)
    # the reference arithmetic: bit-identical to the reference's executor
    out["fp64_err_vs_cpu_reference_arith"] = float(np.max(np.abs(x_refarith.cpu().numpy() - ref)) / den)
    out["reference_arith_bit_identical"] = bool(np.array_equal(x_refarith.cpu().numpy(), ref))
    out["refined_err_vs_cpu_fp64"] = float(np.max(np.abs(xr.cpu().numpy() - ref)) / den)
    out["fp32_err_vs_cpu_fp64"] = float(np.max(np.abs(x32.double().cpu().numpy() - ref)) / den)
    plan.close()
    plan32.close()
    del d_rp, d_ci, d_v, b, x, xr, x32, b32, x_refarith
    if not args.quick:
        out["extra_workloads"] = extra_workloads(args)
    print(json.dumps(out), flush=True)


def _cpu_serial_ms(rp, ci, v, b, seconds):
    """Median time of the reference's serial executor (or the oracle's port of
    it when the reference build is absent) on one host core."""
    import numpy as np

    import oracle
    n, nnz = len(rp) - 1, len(ci)
    cr, cp, cv = oracle.transpose(n, n, nnz, rp, ci, v)
    indeg = oracle.csc_analyser(cr, n, nnz)
    ref = oracle.ref_available("d64")
    run = oracle.ref_csc_executor if ref else oracle.csc_executor
    times, start = [], time.perf_counter()
    run(cp, cr, cv, indeg, n, n, 0, 1, b)
    while time.perf_counter() - start < seconds and len(times) < 1000:
        t = time.perf_counter()
        run(cp, cr, cv, indeg, n, n, 0, 1, b)
        times.append(time.perf_counter() - t)
    return ("reference" if ref else "port"), float(np.median(times)) * 1e3, len(times)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "host CPU"


def _time_gpu(fn, steps, warmup=3):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def extra_workloads(args):
    """The other benchmark workloads, measured the same way (steps back-to-back)."""
    import numpy as np
    import torch

    import oracle
    from capellinisptrsv_amd import Batch, Plan, api

    dev = torch.device("cuda:0")
    res = {}

    # band (n = 100,000, 3 off-diagonals in the previous 1,000 rows, b = L*1) and
    # deep band (n = 500,000, bandwidth 128, p = 0.3): autotuned plans (the
    # deep band resolves to the partitioned solve: within rounding, not
    # bit-exact, so the relative error is reported too)
    for name, cfg in (("band_100k", 0), ("deep_band_500k", 2)):
        log(f"extra workload {name}")
        rp, ci, v = api.make_config(cfg, seed=0)
        n, nnz = len(rp) - 1, len(ci)
        d = [torch.from_numpy(a).to(dev) for a in (rp, ci, v)]
        if cfg == 0:
            b = torch.from_numpy(api.get_x_b(n, n, rp, ci, v)[1]).to(dev)
        else:
            b = torch.from_numpy(api.uniform(1, n)).to(dev)
        x = torch.empty_like(b)
        p = Plan(*d)
        ms = _time_gpu(lambda: p.solve(b, out=x), 20, warmup=2)
        assert p.status() == 0
        ref = oracle.csr_forward(rp, ci, v, b.cpu().numpy(), fused=True)
        err = float(np.max(np.abs(x.cpu().numpy() - ref)))
        levels = api.matrix_layer(n, n, nnz, rp, ci, v)[0]
        ds = api.datasize(n, n, nnz, 8)
        res[name] = {"n": n, "nnzL": nnz, "levels": levels, "plan": p.info(),
                     "ms_per_solve": round(ms, 4), "gflops": round(2 * nnz / (ms * 1e6), 3),
                     "us_per_level": round(ms * 1e3 / levels, 3),
                     "datasize_gb_per_s": round(ds / (ms * 1e6), 2), "frac_hbm": round(ds / (ms * 1e6) / 8000.0, 5),
                     "max_abs_err_vs_oracle": err, "rel_err_vs_oracle": err / float(np.max(np.abs(ref)))}
        if cfg == 0:
            # the thread-per-row choices it competes with, for context
            for algo, kw in (("mix_row_tiles", dict(algo="mix", row_tiles=True)), ("wrtfst", dict(algo="wrtfst"))):
                q = Plan(*d, **kw)
                res[name][f"{algo}_ms_per_solve"] = round(_time_gpu(lambda: q.solve(b, out=x), 20, warmup=2), 4)
                q.close()
        if not args.no_cpu:
            kind, cpu_ms, reps = _cpu_serial_ms(rp, ci, v, b.cpu().numpy(), seconds=3.0)
            res[name]["cpu_serial"] = {"ms_per_solve": round(cpu_ms, 3), "kind": kind, "cores": 1, "solves": reps,
                                       "gpu_speedup": round(cpu_ms / ms, 2)}
        if cfg == 2 and p.info()["algo"] == "partition":
            # 16 right-hand sides at once through the same plan (interleaved)
            k = 16
            bh = api.uniform(5, n * k)
            bk = torch.from_numpy(bh.reshape(n, k)).to(dev)
            xk = torch.empty_like(bk)
            ms_k = _time_gpu(lambda: p.solve(bk, out=xk), 10, warmup=2)
            assert p.status() == 0
            ref = oracle.csr_forward(rp, ci, v, bh, rhs=k, fused=True).reshape(n, k)
            err = float(np.max(np.abs(xk.cpu().numpy() - ref)))
            res[name]["rhs16"] = {"ms_per_solve": round(ms_k, 4), "ms_per_rhs": round(ms_k / k, 4),
                                  "rel_err_vs_oracle": err / float(np.max(np.abs(ref))),
                                  "rhs_block": p.rhs_block(k)}
        p.close()
        if cfg == 2:
            # the same matrix and b in FP32 (autotuned plan), against the FP32 oracle
            v32 = v.astype(np.float32)
            b32 = b.to(torch.float32)
            x32 = torch.empty_like(b32)
            p32 = Plan(d[0], d[1], torch.from_numpy(v32).to(dev))
            ms32 = _time_gpu(lambda: p32.solve(b32, out=x32), 20, warmup=2)
            assert p32.status() == 0
            ref32 = oracle.csr_forward(rp, ci, v32, b32.cpu().numpy(), fused=True)
            err32 = float(np.max(np.abs(x32.cpu().numpy().astype(np.float64) - ref32)))
            res[name]["fp32"] = {"algo": p32.info()["algo"], "ms_per_solve": round(ms32, 4),
                                 "rel_err_vs_fp32_oracle": err32 / float(np.max(np.abs(ref32)))}
            p32.close()

    # 16 right-hand sides on the wide-level matrix: the autotuned plan picks
    # its RHS block (workgroups per CU, threads per workgroup) on the first
    # 16-RHS solve; the fixed plan keeps one 256-thread workgroup per CU
    log("extra workload wide_level_2m_rhs16")
    rp, ci, v = api.make_config(1, seed=0)
    n, nnz, k = len(rp) - 1, len(ci), 16
    d = [torch.from_numpy(a).to(dev) for a in (rp, ci, v)]
    bh = api.uniform(6, n * k)
    bk = torch.from_numpy(bh.reshape(n, k)).to(dev)
    xk = torch.empty_like(bk)
    ref = oracle.csr_forward(rp, ci, v, bh, rhs=k, fused=True).reshape(n, k)
    r = {"n": n, "nnzL": nnz, "k": k}
    for name, kw in (("tuned", {}), ("fixed", dict(algo="wrtfst", sync="value", blocks_per_cu=1, row_threads=256))):
        p = Plan(*d, **kw)
        ms_k = _time_gpu(lambda: p.solve(bk, out=xk), 10, warmup=2)
        assert p.status() == 0
        r[name] = {"ms_per_solve": round(ms_k, 4), "ms_per_rhs": round(ms_k / k, 4),
                   "gflops": round(2 * nnz * k / (ms_k * 1e6), 2),
                   "bit_exact_vs_oracle": bool(np.array_equal(xk.cpu().numpy(), ref)),
                   "rhs_block": p.rhs_block(k), "plan": p.info()["algo"]}
        p.close()
    res["wide_level_2m_rhs16"] = r

    # differentiable solve: n = 1,000,000 level matrix, forward + adjoint
    log("extra workload adjoint_1m")
    rp, ci, v = api.make_config(4, seed=0)
    n, nnz = len(rp) - 1, len(ci)
    d = [torch.from_numpy(a).to(dev) for a in (rp, ci, v)]
    b = torch.from_numpy(api.uniform(1, n)).to(dev)
    xt = torch.from_numpy(api.uniform(2, n)).to(dev)
    r = {"n": n, "nnzL": nnz}
    for mode in ("gather", "scatter"):
        p = Plan(*d, algo="wrtfst", sync="value", adjoint=mode)
        x = p.solve(b)
        g = (x - xt).contiguous()
        fwd = _time_gpu(lambda: p.solve(b, out=x), 20)
        adj = _time_gpu(lambda: p.adjoint(g, x), 20)
        assert p.status() == 0
        r[mode] = {"forward_ms": round(fwd, 4), "adjoint_ms": round(adj, 4), "adjoint_over_forward": round(adj / fwd, 3)}
        p.close()
    res["adjoint_1m"] = r

    # batched: 256 matrices, n log-uniform in [2,000, 200,000], nnz/row U{1..16}
    # (diagonal included), band U[64, 4096], k = 16 interleaved RHS
    log("extra workload batched_256x16")
    mats, row_off, rp_all, ci_all, v_all = api.make_batch(256, seed=0)
    k = 16
    N = int(row_off[-1])
    bb = api.uniform(3, N * k)
    t = [torch.from_numpy(a).to(dev) for a in (row_off, rp_all, ci_all, v_all)]
    bt = Batch(*t, k=k)
    bd = torch.from_numpy(bb.reshape(N, k)).to(dev)
    xb = torch.empty_like(bd)
    fp64_ms = _time_gpu(lambda: bt.solve(bd, out=xb), 5, warmup=1)
    x64 = xb.clone()
    xr = torch.empty_like(bd)
    holder = {}

    def refined():
        holder["r"] = bt.solve_refined(bd, tol=1e-12, out=xr)
    ref_ms = _time_gpu(refined, 3, warmup=1)
    _, iters, rel = holder["r"]
    # error against the FP64 oracle, every matrix
    worst64, worstr = 0.0, 0.0
    x64h, xrh = x64.reshape(-1).cpu().numpy(), xr.reshape(-1).cpu().numpy()
    for m in range(256):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        o = oracle.csr_forward(*mats[m], bb[lo:hi], rhs=k, fused=True)
        den = float(np.max(np.abs(o)))
        worst64 = max(worst64, float(np.max(np.abs(x64h[lo:hi] - o))) / den)
        worstr = max(worstr, float(np.max(np.abs(xrh[lo:hi] - o))) / den)
    res["batched_256x16"] = {
        "matrices": 256, "rows": N, "nnzL": int(len(ci_all)), "k": k,
        "pick": {"algo": bt.algo, "blocks_per_cu": bt.blocks_per_cu, "row_threads": bt.row_threads,
                 "rhs_per_lane": bt.rhs_per_lane, "row_tiles": bt.row_tiles,
                 "partitioned_matrices": bt.partitioned_matrices},
        "fp64_ms": round(fp64_ms, 3), "fp64_gflops": round(2 * len(ci_all) * k / (fp64_ms * 1e6), 2),
        "fp32_refined_ms": round(ref_ms, 3), "refined_iterations": iters, "refined_over_fp64": round(ref_ms / fp64_ms, 3),
        "fp64_rel_err_max_over_matrices": worst64, "refined_rel_err_max_over_matrices": worstr,
    }
    bt.close()
    res["rocsparse_csrsv"] = rocsparse_comparison()
    res["algorithms_wide_level_2m"] = algorithm_table()
    return res


def algorithm_table():
    """Every solve algorithm of the library on the wide-level matrix (the
    reference's four drop-in paths and the one-workgroup kernel), each with
    its error against the oracle."""
    import numpy as np
    import torch

    import oracle
    from capellinisptrsv_amd import Plan, api

    log("algorithm table")
    dev = torch.device("cuda:0")
    rp, ci, v = api.make_config(1, seed=0)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(dev) for a in (rp, ci, v)]
    b = torch.from_numpy(api.uniform(1, n)).to(dev)
    x = torch.empty_like(b)
    ref = oracle.csr_forward(rp, ci, v, b.cpu().numpy(), fused=True)
    out = {}
    for name, kw in (("wrtfst_value", dict(algo="wrtfst", sync="value")),
                     ("wrtfst_value_lds", dict(algo="wrtfst", sync="value", lds_window=True)),
                     ("wrtfst_flag", dict(algo="wrtfst", sync="flag")),
                     ("twopart_value", dict(algo="twopart", sync="value")),
                     ("mix_value", dict(algo="mix", sync="value")),
                     ("csc", dict(algo="csc")),
                     ("local", dict(algo="local"))):
        p = Plan(*d, **kw)
        ms = _time_gpu(lambda: p.solve(b, out=x), 5 if name == "local" else 20, warmup=2)
        err = float(np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref)))
        out[name] = {"ms_per_solve": round(ms, 4), "gflops": round(2 * len(ci) / (ms * 1e6), 3),
                     "rel_err_vs_oracle": err, "status": p.status()}
        p.close()
    return out


def rocsparse_comparison():
    """rocSPARSE csrsv (analysis once, then solves) on three of the matrices,
    next to this library's autotuned plan: context for the numbers above."""
    import numpy as np
    import torch

    import oracle
    from capellinisptrsv_amd import Plan, api
    from capellinisptrsv_amd.rocsparse_cmp import Csrsv

    dev = torch.device("cuda:0")
    out = {}
    for name, cfg in (("wide_level_2m", 1), ("band_100k", 0), ("deep_band_500k", 2)):
        log(f"rocsparse comparison {name}")
        rp, ci, v = api.make_config(cfg, seed=0)
        n = len(rp) - 1
        d = [torch.from_numpy(a).to(dev) for a in (rp, ci, v)]
        b = torch.from_numpy(api.uniform(1, n)).to(dev)
        x = torch.empty_like(b)
        r = {}
        try:
            cs = Csrsv(*d)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cs.analysis()
            torch.cuda.synchronize()
            r["rocsparse_analysis_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
            steps = 20 if cfg != 2 else 3
            r["rocsparse_ms_per_solve"] = round(_time_gpu(lambda: cs.solve(b, x), steps, warmup=2), 4)
            ref = oracle.csr_forward(rp, ci, v, b.cpu().numpy(), fused=True)
            r["rocsparse_rel_err_vs_oracle"] = float(np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref)))
            cs.close()
        except Exception as e:  # a comparator that cannot run is reported, not fatal
            r["rocsparse_error"] = f"{type(e).__name__}: {e}"
        p = Plan(*d)
        r["this_library_ms_per_solve"] = round(_time_gpu(lambda: p.solve(b, out=x), 20 if cfg != 2 else 3, warmup=2), 4)
        r["this_library_algo"] = p.info()["algo"]
        if "rocsparse_ms_per_solve" in r:
            r["speedup_over_rocsparse"] = round(r["rocsparse_ms_per_solve"] / r["this_library_ms_per_solve"], 2)
        p.close()
        out[name] = r
    return out


if __name__ == "__main__":
    main()
