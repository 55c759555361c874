"""The stage pair's LDS ring protocol, as a model run on the CPU.

The two waves of a stage-pair tile (dim2d_kernels.hip: pair_producer,
pair_consumer, pair_tile) share an 8-slot ring of stage-1 rows and pass the
same number of barriers, kPairLag + (r1 - r0 + 2): the producer writes one row
per barrier interval (and then the north wall ghost rows when its tile holds
them), the consumer idles kPairLag intervals, then (with the south wall ghost
rows first when its tile gathers them) reads one new row per interval.  This
model replays the operations of both waves interval by interval, the two
waves' operations within an interval in random order, for every edge kind of
the pair (reflection, periodic, S/N walls, open), grid heights down to the
smallest the pair takes and tile heights from one row to the whole grid.  It
checks that both waves pass the same number of barriers and that every ring
read sees the row, and the writer (stage 1 or a wall step), it should: no slot
is reused while a row in it is still to be read, and no row is read before it
is written.
"""
import random

import pytest

KBC = 2     # ghost layers
SLOTS = 8   # ring rows
LAG = 6     # idle intervals of the consumer


def gather_row(ny, bc_s, bc_n, x):
    """The ring row pair_gather reads for row x (None: a frame cell from HBM)."""
    if bc_s == "open" and (x < KBC or x >= ny - KBC):
        return None
    if x < KBC and bc_s == "reflection":
        return 2 * KBC - 1 - x
    if x >= ny - KBC and bc_n == "reflection":
        return 2 * ny - 2 * KBC - 1 - x
    return x


def writer_of(ny, bc_s, bc_n, row):
    if (row < KBC and bc_s == "wall") or (row >= ny - KBC and bc_n == "wall"):
        return "wall"
    return "stage1"


def producer(ny, bc_s, bc_n, r0, r1):
    """The producer's operations, one list per barrier interval."""
    rlast = r1 + 1 if bc_n == "periodic" else min(r1 + 1, ny - KBC - 1)
    top = r1 + 1 >= ny - KBC
    phases = []
    for r in range(r0 - 4, rlast + 1):
        phases.append([("write", r, "stage1")] if r >= r0 - 2 else [])
    if top and bc_n == "wall":
        phases.append([("read_s", ny - 3, "stage1"), ("read_s", ny - 4, "stage1"),
                       ("write_s", ny - 2, "wall"), ("write_s", ny - 1, "wall")])
    return phases


def consumer(ny, bc_s, bc_n, r0, r1):
    phases = [[] for _ in range(LAG)]

    def gather(x):
        rr = gather_row(ny, bc_s, bc_n, x)
        return [] if rr is None else [("read_s", rr, writer_of(ny, bc_s, bc_n, rr))]

    first = []
    if r0 - 2 < KBC and bc_s == "wall":
        first += [("read_s", 2, "stage1"), ("read_s", 3, "stage1"), ("write_s", 1, "wall")]
        if r0 == KBC:
            first.append(("write_s", 0, "wall"))
    first += gather(r0 - 2) + gather(r0 - 1)
    for r in range(r0 - 2, r1):
        ops = first + gather(r + 2)
        first = []
        if r >= r0:
            ops.append(("read_u", r, "stage1"))
        phases.append(ops)
    return phases


def run_tile(ny, bc_s, bc_n, r0, r1, rng):
    steps = LAG + (r1 - r0 + 2)
    prod, cons = producer(ny, bc_s, bc_n, r0, r1), consumer(ny, bc_s, bc_n, r0, r1)
    where = f"ny={ny} {bc_s}/{bc_n} tile [{r0},{r1})"
    assert len(cons) == steps, where
    assert len(prod) <= steps, f"{where}: the producer's work needs more barriers than the tile has"
    prod += [[] for _ in range(steps - len(prod))]
    s_slot = [None] * SLOTS
    u_slot = [None] * SLOTS
    for k in range(steps):
        a, b = list(prod[k]), list(cons[k])
        while a or b:
            take_a = a and (not b or rng.random() < 0.5)
            kind, row, who = a.pop(0) if take_a else b.pop(0)
            if kind == "write":
                s_slot[row % SLOTS] = (row, who)
                u_slot[row % SLOTS] = (row, who)
            elif kind == "write_s":
                s_slot[row % SLOTS] = (row, who)
            elif kind == "read_s":
                got = s_slot[row % SLOTS]
                assert got == (row, who), f"{where}: s row {row} holds {got} (interval {k})"
            elif kind == "read_u":
                got = u_slot[row % SLOTS]
                assert got == (row, who), f"{where}: u row {row} holds {got} (interval {k})"


EDGES = [("reflection", "reflection"), ("periodic", "periodic"), ("wall", "wall"),
         ("wall", "reflection"), ("reflection", "wall"), ("open", "open")]


@pytest.mark.parametrize("bc_s,bc_n", EDGES)
def test_pair_ring_reads_every_row_as_written(bc_s, bc_n):
    rng = random.Random(7)
    for ny in (12, 13, 17, 30, 41):
        nj = ny - 2 * KBC
        for rows in (1, 2, 3, 4, 7, 8, 32, nj):
            for r0 in range(KBC, ny - KBC, rows):
                r1 = min(r0 + rows, ny - KBC)
                for _ in range(4):
                    run_tile(ny, bc_s, bc_n, r0, r1, rng)


def test_the_model_sees_a_missing_wall_step():
    """A consumer that computed the south wall rows only in the tile starting
    at row 2 would, in one-row tiles, read the stage-1 value of ghost row 1:
    the model must see that."""
    global consumer
    saved = consumer

    def only_first_tile(ny, bc_s, bc_n, r0, r1):
        phases = saved(ny, bc_s, bc_n, r0, r1)
        if r0 != KBC:
            phases = [[op for op in ph if op[0] != "write_s" and op[1:] not in
                       ((2, "stage1"), (3, "stage1"))] for ph in phases]
        return phases

    consumer = only_first_tile
    try:
        with pytest.raises(AssertionError):
            run_tile(30, "wall", "wall", 3, 4, random.Random(1))
    finally:
        consumer = saved


def test_the_model_sees_a_short_lag():
    """Interior rows would do with one idle interval fewer, but the first
    interval of a bottom tile at a south wall (or a reflecting south edge)
    reads row r0 + 1, which the producer writes in that same interval with
    five idle intervals: some interleaving then reads it before it is
    written."""
    global LAG
    saved = LAG
    LAG = saved - 1
    try:
        for bc_s in ("wall", "reflection"):
            with pytest.raises(AssertionError):
                for seed in range(40):
                    run_tile(30, bc_s, "reflection", 2, 14, random.Random(seed))
    finally:
        LAG = saved
