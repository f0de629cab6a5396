"""Index arithmetic of k_mg_up0's LDS window, simulated on the CPU.

k_mg_up0 (fg_mg.hip) keeps each fine cell's smoothing start in a ring of
RB = 2J + 2P floats, slot f mod RB, P = 1024 cells a pass (four a thread).
Pass p reads the cell and its neighbours (2-D clamped) for cells p .. p+P-1
of the chunk; the prologue fills [begin - J, begin + J + P), and each pass
writes the P cells the next pass adds, [p + J + P, p + J + 2P).  This
replays that schedule for many grid and chunk shapes and checks that every
read finds the cell it wants and that no write lands on a cell the same
pass reads.  (The same schedule with flat +-1, +-J neighbours was checked
for the direction update's window, which is not used.)
"""
import random

import pytest

BLOCK, ITEMS = 256, 4
PASS = BLOCK * ITEMS


def _run(I, J, begin, end):
    C = I * J
    RB = 2 * J + 2 * PASS
    ring = [None] * RB
    for f in range(max(begin - J, 0), min(begin + J + PASS, C)):
        ring[f % RB] = f
    sp, p = begin % RB, begin
    while p < end:
        writes = []
        for t in range(BLOCK):
            for u in range(ITEMS):
                f = p + J + PASS + t + u * BLOCK
                if p + PASS < end and f < C:
                    writes.append(f)
                c = p + t + u * BLOCK
                if c >= end:
                    continue
                ci, cj = divmod(c, J)
                sc = sp + t + u * BLOCK
                sc -= RB if sc >= RB else 0
                for di, dj in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
                    gi, gj = min(max(ci + di, 0), I - 1), min(max(cj + dj, 0), J - 1)
                    sg = sc + (gi - ci) * J + (gj - cj)
                    sg += RB if sg < 0 else 0
                    sg -= RB if sg >= RB else 0
                    assert ring[sg] == gi * J + gj, (I, J, begin, end, p, c)
        for f in writes:
            old = ring[f % RB]
            assert old is None or old < p - J, (old, p, J)
            ring[f % RB] = f
        p += PASS
        sp += PASS
        sp -= RB if sp >= RB else 0


def _cases():
    fixed = [(64, 64, 0, 1024), (64, 64, 3072, 4096), (100, 37, 1024, 2048), (100, 37, 3584, 3700),
             (5, 3, 0, 15), (257, 300, 76800, 77100), (64, 4096, 16384, 32768), (40, 1024, 36864, 40960)]
    rng = random.Random(3)
    for _ in range(12):
        I, J = rng.randint(2, 200), rng.randint(1, 400)
        chunk = rng.choice([256, 1024, 4096])
        b = rng.randrange((I * J + chunk - 1) // chunk)
        fixed.append((I, J, b * chunk, min(I * J, (b + 1) * chunk)))
    return fixed


@pytest.mark.parametrize("I,J,begin,end", _cases())
def test_window_reads_find_their_cells(I, J, begin, end):
    _run(I, J, begin, end)
