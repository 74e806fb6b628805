"""CPU lockstep simulation of the write-first schedule (forward progress, no GPU).

Models the kernel's execution: workgroups take row tickets in the order they
start; the 64 lanes of a wavefront execute the polling loop in lockstep
(every load of an iteration happens before its stores); wavefronts of the
device interleave in a random order; at most `resident` wavefronts run at
once.  Two loop shapes are simulated:

  uniform  -- the shipped shape: a lane publishes inside the loop, the loop
              exits when every lane of the wavefront has published;
  sunk     -- the shape the compiler produced from a per-lane loop exit: the
              publishing store executes only after every lane has left the
              loop.

The uniform shape must always finish with the oracle's values; the sunk
shape must deadlock as soon as a row depends on another row of its own
wavefront (the bug the library's kernels are structured against).
"""
import numpy as np
import pytest

import oracle
from conftest import random_lower

SENT = object()


class Wave:
    def __init__(self, rows, rp, ci, v, b, sunk):
        self.rows = rows
        self.rp, self.ci, self.v, self.b = rp, ci, v, b
        self.j = {i: rp[i] for i in rows}
        self.s = {i: 0.0 for i in rows}
        self.done = {i: False for i in rows}
        self.pending = {}  # sunk shape: values computed but not yet stored
        self.sunk = sunk

    def step(self, x):
        """One lockstep iteration; returns True if any lane made progress."""
        reads = {}
        for i in self.rows:  # loads first (the whole wavefront)
            if not self.done[i] and self.j[i] < self.rp[i + 1] - 1:
                reads[i] = x[self.ci[self.j[i]]]
        progress = False
        stores = {}
        for i in self.rows:
            if self.done[i]:
                continue
            d = self.rp[i + 1] - 1
            if i in reads and reads[i] is not SENT:
                p = self.v[self.j[i]] * reads[i]
                self.s[i] = self.s[i] + p
                self.j[i] += 1
                progress = True
            if self.j[i] >= d:
                xi = (self.b[i] - self.s[i]) / self.v[d]
                self.done[i] = True
                progress = True
                if self.sunk:
                    self.pending[i] = xi
                else:
                    stores[i] = xi
        if self.sunk and all(self.done.values()):
            stores.update(self.pending)
            self.pending = {}
        for i, val in stores.items():
            x[i] = val
        return progress

    def finished(self):
        return all(self.done.values()) and not self.pending


def simulate(rp, ci, v, b, sunk=False, resident=6, rows_per_ticket=64, seed=0, max_idle=20000):
    rng = np.random.default_rng(seed)
    n = len(rp) - 1
    x = [SENT] * n
    next_row = 0
    running = []
    idle = 0
    while next_row < n or running:
        while len(running) < resident and next_row < n:  # a resident slot takes the next ticket
            rows = list(range(next_row, min(n, next_row + rows_per_ticket)))
            running.append(Wave(rows, rp, ci, v, b, sunk))
            next_row += rows_per_ticket
        w = running[int(rng.integers(len(running)))]
        idle = 0 if w.step(x) else idle + 1
        if w.finished():
            running.remove(w)
        if idle > max_idle:
            return None  # no wavefront can progress: deadlock
    return np.array(x, dtype=np.float64)


@pytest.mark.parametrize("seed", range(4))
def test_uniform_exit_always_completes(seed):
    rp, ci, v = random_lower(700, max_off=4, band=90, seed=seed)
    b = np.random.default_rng(seed).uniform(-1, 1, size=700)
    x = simulate(rp, ci, v, b, sunk=False, resident=[1, 2, 5, 11][seed], seed=seed)
    assert x is not None
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_uniform_exit_chain_single_resident():
    # a bidiagonal chain: every row depends on the previous lane of its wavefront
    n = 300
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[1.0, np.tile([0.5, 2.0], n - 1)]
    b = np.linspace(-1, 1, n)
    x = simulate(rp, ci, v, b, resident=1)
    assert x is not None and np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_sunk_store_deadlocks_on_intra_wavefront_dependency():
    n = 300
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[1.0, np.tile([0.5, 2.0], n - 1)]
    b = np.linspace(-1, 1, n)
    assert simulate(rp, ci, v, b, sunk=True, resident=4, max_idle=2000) is None


def test_sunk_store_survives_without_intra_wavefront_dependencies():
    # level structure with every dependency >= one ticket back: no deadlock even sunk
    rp, ci, v = random_lower(640, max_off=3, band=64, seed=3)
    keep_rp, keep_ci, keep_v = [0], [], []
    for i in range(640):
        cols = [int(c) for c in ci[rp[i]:rp[i + 1] - 1] if c < (i // 64) * 64]
        vals = [float(vv) for c, vv in zip(ci[rp[i]:rp[i + 1] - 1], v[rp[i]:rp[i + 1] - 1]) if c < (i // 64) * 64]
        keep_ci += cols + [i]
        keep_v += vals + [1.0 + sum(abs(t) for t in vals)]
        keep_rp.append(len(keep_ci))
    rp2, ci2, v2 = np.array(keep_rp, np.int32), np.array(keep_ci, np.int32), np.array(keep_v)
    b = np.ones(640)
    x = simulate(rp2, ci2, v2, b, sunk=True, resident=3)
    assert x is not None and np.array_equal(x, oracle.csr_forward(rp2, ci2, v2, b, fused=False))


# ---------------------------------------------------------------- batches
# The batch launch: the matrices are concatenated (block-diagonal, global
# columns); blocks of `rows` consecutive rows of one matrix are handed out by
# ticket in the order of their relative level -- the running maximum level at
# the block over the matrix's depth, stable-sorted (the rule of
# build_batch_schedule) -- to workgroups of `waves` wavefronts.  A workgroup
# splits its block over its wavefronts (64 rows each), waits at a barrier
# until all of them have published, then takes its next ticket.  At most
# `resident` workgroups run at once; the others are admitted one by one only
# when a resident workgroup has left (its tickets ran out), the way a
# persistent grid larger than the device is scheduled.

def batch_schedule(row_off, lv, rows):
    blks = []
    for m in range(len(row_off) - 1):
        lo_m, hi_m = int(row_off[m]), int(row_off[m + 1])
        if hi_m == lo_m:
            continue
        depth = max(int(lv[r]) + 1 for r in range(lo_m, hi_m))
        run = 0
        for lo in range(lo_m, hi_m, rows):
            hi = min(lo + rows, hi_m)
            run = max(run, int(lv[lo]))
            blks.append((run / depth, lo, hi - lo))
            run = max([run] + [int(lv[r]) for r in range(lo, hi)])
    blks.sort(key=lambda t: t[0])  # stable, as std::stable_sort
    return [(lo, ln) for _, lo, ln in blks]


class Workgroup:
    def __init__(self, waves):
        self.waves = waves
        self.block = None
        self.members = []

    def idle(self):
        return all(w.finished() for w in self.members)


def simulate_batch(rp, ci, v, b, sched, waves=4, resident=3, n_wg=6, seed=0, max_idle=50000):
    rng = np.random.default_rng(seed)
    n = len(rp) - 1
    x = [SENT] * n
    next_ticket = 0
    pending = list(range(n_wg))  # workgroups not yet admitted
    running = []
    idle = 0
    while pending or running:
        while len(running) < resident and pending:  # late admission
            pending.pop(0)
            running.append(Workgroup(waves))
        wg = running[int(rng.integers(len(running)))]
        if wg.idle():  # barrier passed: the next ticket
            if next_ticket >= len(sched):
                running.remove(wg)
                continue
            lo, ln = sched[next_ticket]
            next_ticket += 1
            wg.members = [Wave(list(range(lo + 64 * w, min(lo + ln, lo + 64 * (w + 1)))), rp, ci, v, b, False)
                          for w in range(waves) if 64 * w < ln]
            idle = 0
            continue
        live = [w for w in wg.members if not w.finished()]
        w = live[int(rng.integers(len(live)))]
        idle = 0 if w.step(x) else idle + 1
        if idle > max_idle:
            return None
    return np.array(x, dtype=np.float64)


def _batch_of(seed, n_mat):
    rng = np.random.default_rng(seed)
    mats = []
    for m in range(n_mat):
        nm = int(rng.integers(40, 400))
        band = int(rng.integers(2, 120))
        mats.append(random_lower(nm, max_off=int(rng.integers(1, 6)), band=band, seed=seed * 100 + m))
    # a chain: one row per level, the deepest member
    c = 300
    mats.append((np.r_[0, 1, np.arange(3, 2 * c, 2)].astype(np.int32),
                 np.r_[0, np.c_[np.arange(0, c - 1), np.arange(1, c)].reshape(-1)].astype(np.int32),
                 np.r_[1.0, np.tile([0.5, 2.0], c - 1)]))
    row_off = np.cumsum([0] + [len(m[0]) - 1 for m in mats])
    nnz_off = np.cumsum([0] + [len(m[1]) for m in mats])
    rp = np.concatenate([m[0][:-1] + o for m, o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    ci = np.concatenate([m[1] + ro for m, ro in zip(mats, row_off)]).astype(np.int32)  # global columns
    v = np.concatenate([m[2] for m in mats])
    return rp, ci, v, row_off


@pytest.mark.parametrize("seed,rows,waves,resident,n_wg", [(0, 256, 4, 2, 8), (1, 128, 2, 1, 4), (2, 256, 4, 5, 5),
                                                           (3, 64, 1, 3, 12), (4, 256, 4, 8, 64)])
def test_batch_relative_level_schedule_always_completes(seed, rows, waves, resident, n_wg):
    """Tickets in relative-level order, workgroup barriers between tickets and
    workgroups admitted late (fewer resident than launched, up to eight times
    as many launched as resident): the batch always completes with the
    oracle's values, whatever the interleaving."""
    rp, ci, v, row_off = _batch_of(seed, 6)
    n = len(rp) - 1
    b = np.random.default_rng(seed).uniform(-1, 1, size=n)
    lv = oracle.row_levels(n, rp, ci)
    sched = batch_schedule(row_off, lv, rows)
    # every matrix keeps its row order inside the schedule
    for m in range(len(row_off) - 1):
        own = [lo for lo, _ in sched if row_off[m] <= lo < row_off[m + 1]]
        assert own == sorted(own)
    x = simulate_batch(rp, ci, v, b, sched, waves=waves, resident=resident, n_wg=n_wg, seed=seed)
    assert x is not None
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_batch_schedule_with_a_block_order_broken_deadlocks():
    """The check has teeth: hand a matrix's blocks out in reverse order with a
    single resident workgroup and the simulation detects the deadlock."""
    rp, ci, v, row_off = _batch_of(5, 2)
    n = len(rp) - 1
    b = np.ones(n)
    lv = oracle.row_levels(n, rp, ci)
    sched = batch_schedule(row_off, lv, 64)[::-1]
    assert simulate_batch(rp, ci, v, b, sched, waves=1, resident=1, n_wg=1, seed=0, max_idle=3000) is None
