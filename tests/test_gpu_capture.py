"""Solves captured into a graph (torch.cuda.CUDAGraph over the caller's
stream) and solves on a caller's non-default stream.

A plan's solve and a hybrid batch's solve (its partitioned matrices on the
batch's own side streams, forked from and joined back into the caller's
stream) must be capturable after one eager solve with the same shapes, and
every replay -- with eager work and synchronizations between replays -- must
give the eager result bit for bit.  On a non-default stream, work the caller
queues right after a batch solve must see the whole result (the join).
"""
import numpy as np
import pytest

from capellinisptrsv_amd import api

pytestmark = pytest.mark.gpu


def _replays(run, out, eager, n=3):
    import torch
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        run()
    for rep in range(n):
        out.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep


def test_plan_solve_replays(cuda):
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(1, seed=0, scale=0.05)
    n = len(rp) - 1
    plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), algo="wrtfst", sync="value")
    b = torch.from_numpy(api.uniform(61, n)).to(cuda)
    eager = plan.solve(b).clone()
    out = torch.empty_like(b)
    _replays(lambda: plan.solve(b, out=out, stream=torch.cuda.current_stream()), out, eager)
    assert plan.status() == 0
    plan.close()


def test_hybrid_batch_solve_replays_and_user_stream(cuda):
    import torch
    from capellinisptrsv_amd import Batch
    from test_gpu_refine_batch import _hybrid_batch
    mats, row_off, rp, ci, v = _hybrid_batch()
    k = 4
    N = int(row_off[-1])
    bt = Batch(*(torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)), k=k, algo="auto")
    assert bt.partitioned >= 2  # several side streams in use
    b = torch.from_numpy(api.uniform(62, N * k).reshape(N, k)).to(cuda)
    eager = bt.solve(b).clone()
    torch.cuda.synchronize()
    # a caller's stream: the copy queued right behind the solve sees every row
    s = torch.cuda.Stream()
    x = torch.empty_like(b)
    y = torch.empty_like(b)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            x.fill_(-1.0)
            bt.solve(b, out=x, stream=s)
            y.copy_(x)
    torch.cuda.synchronize()
    assert torch.equal(y, eager)
    # captured, then replayed
    out = torch.empty_like(b)
    _replays(lambda: bt.solve(b, out=out, stream=torch.cuda.current_stream()), out, eager)
    bt.close()
