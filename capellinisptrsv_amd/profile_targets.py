"""Small GPU workloads for rocprofv3 (kernel traces and counter runs).

    python -m capellinisptrsv_amd.profile_targets solve [--steps N] [--algo auto]
        the wide-level n = 2,000,000 FP64 solve of bench.py, N reset+solve steps
    python -m capellinisptrsv_amd.profile_targets copy [--mib 1024]
        out = in * 1 over a float64 buffer of the given size: a kernel that
        reads and writes exactly mib MiB each, to calibrate FETCH_SIZE /
        WRITE_SIZE against known bytes
    python -m capellinisptrsv_amd.profile_targets adjoint [--steps N]
        forward + gather adjoint of the n = 1,000,000 level matrix
    python -m capellinisptrsv_amd.profile_targets batched [--steps N]
        the 256-matrix batch with 16 right-hand sides: N FP64 solves and N
        FP32 + FP64 refined solves
    python -m capellinisptrsv_amd.profile_targets refine [--steps N]
        the wide-level matrix: N FP64 solves, then N FP32 + FP64 refined
        solves (fused first two steps) to 1e-12
    python -m capellinisptrsv_amd.profile_targets band [--steps N] [--algo auto]
        the n = 100,000 band (3 off-diagonals in the previous 1,000 rows,
        723 levels), N reset+solve steps
    python -m capellinisptrsv_amd.profile_targets deep [--steps N] [--algo auto] [--rhs K]
        the n = 500,000 bandwidth-128 matrix (~38 nnz/row, 197,032 levels);
        --algo partition builds the partitioned solve without the autotuner;
        --rhs K solves K right-hand sides at once
"""
import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("target", choices=["solve", "copy", "adjoint", "batched", "deep", "refine", "band"])
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--algo", default="auto")
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--rhs", type=int, default=1)
    a = ap.parse_args(argv)

    import torch

    from capellinisptrsv_amd import Plan, api

    dev = torch.device("cuda:0")
    if a.target == "copy":
        n = a.mib * 1024 * 1024 // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(a.steps):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(f"copy: {a.steps} x {a.mib} MiB read + {a.mib} MiB written", flush=True)
        return 0
    if a.target == "batched":
        from capellinisptrsv_amd import Batch
        _, row_off, rp, ci, v = api.make_batch(256, seed=0)
        k, n = 16, int(row_off[-1])
        bt = Batch(*(torch.from_numpy(t).to(dev) for t in (row_off, rp, ci, v)), k=k)
        print(f"batched: {bt.algo}, {bt.blocks_per_cu} workgroup(s) per CU, {bt.row_threads} threads, "
              f"{bt.rhs_per_lane} RHS per lane, partitioned matrices {bt.partitioned_matrices}", flush=True)
        b = torch.from_numpy(api.uniform(3, n * k).reshape(n, k)).to(dev)
        x = torch.empty_like(b)
        for _ in range(a.steps):
            bt.solve(b, out=x)
        for _ in range(a.steps):
            _, it, rel = bt.solve_refined(b, tol=1e-12, out=x)
        torch.cuda.synchronize()
        print(f"batched: n={n} nnz={len(ci)} k={k} steps={a.steps} refined iterations={it}", flush=True)
        bt.close()
        return 0
    cfg = {"solve": 1, "refine": 1, "adjoint": 4, "deep": 2, "band": 0}[a.target]
    rp, ci, v = api.make_config(cfg, seed=0)
    n = len(rp) - 1
    d = [torch.from_numpy(x).to(dev) for x in (rp, ci, v)]
    b = torch.from_numpy(api.uniform(1, n)).to(dev)
    if a.target in ("deep", "solve") and a.rhs > 1:
        b = torch.from_numpy(api.uniform(1, n * a.rhs).reshape(n, a.rhs)).to(dev)
    x = torch.empty_like(b)
    plan = Plan(*d, algo=a.algo) if a.target in ("solve", "deep", "refine", "band") else \
        Plan(*d, algo="wrtfst", sync="value")
    for _ in range(a.steps):
        plan.solve(b, out=x)
        if a.target == "adjoint":
            plan.adjoint(x, x)
    if a.target == "refine":
        xr = torch.empty_like(b)
        for _ in range(a.steps):
            _, it, rel = plan.solve_refined(b, tol=1e-12, out=xr)
        print(f"refined: iterations={it} last correction={rel:.3e}", flush=True)
    torch.cuda.synchronize()
    assert plan.status() == 0
    pick = plan.rhs_block(a.rhs) if a.rhs > 1 else None
    print(f"{a.target}: n={n} nnz={len(ci)} plan={plan.info()} rhs={a.rhs} rhs block={pick} steps={a.steps}",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
