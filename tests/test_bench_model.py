"""The closed-form critical path of the temporal-blocked solve (bench.py's latency bound) against
a block-by-block schedule of the kernel's own rules, with every tile running as early as they
allow and one time unit per column step:
- before block b a tile stages block b + 1, so it waits until its row-tile producer (g - 1, kb)
  has published row s1 + 64 and its sweep-block producer (g, kb - 1) row s1 + TK (clamped to the
  tile's rows), s1 being block b + 1's last step; the prologue waits for block 0's rows;
- a producer publishes rows [0, k BLK) when its block k starts (on time), halfway through block k
  (halfway), or when block k + 1 starts (one block late), and all rows after its last block."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def ceil_div(a, b):
    return -(-a // b)


def scheduled_finish(G, KB, rows, TK, BLK, lag):
    nblk = rows // BLK
    pub = {}

    def ready(g, kb, b):
        s1 = b * BLK + BLK - 1
        t = 0
        if g > 0:
            t = max(t, pub[(g - 1, kb, ceil_div(min(s1 + 64, rows), BLK))])
        if kb > 0:
            t = max(t, pub[(g, kb - 1, ceil_div(min(s1 + TK, rows), BLK))])
        return t

    finish = 0
    for g in range(G):
        for kb in range(KB):
            t = ready(g, kb, 0)
            for b in range(nblk):
                if b > 0 and lag == 1:
                    pub[(g, kb, b)] = t
                if b > 1 and lag == 2:
                    pub[(g, kb, b - 1)] = t
                if b + 1 < nblk:
                    t = max(t, ready(g, kb, b + 1))
                if b > 0 and lag == 3:
                    pub[(g, kb, b)] = t + BLK // 2
                t += BLK
            for k in range(1, nblk + 1):
                pub.setdefault((g, kb, k), t)
            pub[(g, kb, nblk)] = t
            finish = max(finish, t)
    return finish


CASES = [(5, 5, 328, 8, 8, 1), (5, 5, 328, 8, 8, 2), (5, 5, 328, 8, 8, 3), (3, 3, 164, 8, 4, 3),
         (10, 10, 584, 8, 8, 1), (4, 6, 200, 16, 8, 1), (6, 4, 320, 4, 16, 2), (1, 1, 72, 2, 8, 1),
         (68, 25, 4176, 8, 8, 3)]


@pytest.mark.parametrize("case", CASES, ids=["G%d_KB%d_rows%d_tk%d_blk%d_lag%d" % c for c in CASES])
def test_critical_path_formula_matches_the_schedule(case):
    G, KB, rows, TK, BLK, lag = case
    layout = dict(sweeps_per_tile=TK, cols_per_block=BLK, handoff_lag=lag, row_tiles=G, sweep_blocks=KB,
                  steps_per_tile=rows)
    assert load_bench().critical_path_steps(layout) == scheduled_finish(G, KB, rows, TK, BLK, lag)
