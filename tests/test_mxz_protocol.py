"""Model check of the progress-counter protocol that couples the waves of
k_mxz (Maxwell + the next step's sweeps in one launch).  M Maxwell waves share
the ring rows in tasks of one (FP32) or two (FP64) rows dealt round-robin
(wave m, two-row tasks: rows 2m, 2m + 1, 2m + 2M, 2m + 2M + 1, ...); for
a task of rows k .. k' a Maxwell wave waits until every sweep wave has
released row k' - S, writes the slots of its rows, then publishes
ready[k % S] = k + 1 for each row in turn.  A sweep wave at ring row kc
waits for ready[kc % S] >= kc + 1 and ready[(kc - 1) % S] >= kc, reads rows kc and
kc - 1, then releases the rows below the ones its next row reads.  Writing a
slot, publishing it, reading the ring and releasing are separate steps.
Random interleavings of all waves over the row sequences of many chunk shapes
must finish (no wait without end) and never overwrite a row that a sweep wave
may still read, nor let a sweep wave read a slot that holds another row."""
import random

import pytest

S = 8       # MXZ_SLOTS


def run(z0, tz, nz, nmax, seed, task=2):
    rnd = random.Random(seed)
    zend = min(z0 + tz, nz)
    elast = zend + 2
    rc0 = max(0, min(z0 - 3, nz - 1))

    def rowof(e):
        return max(0, min(min(e, elast), nz - 1))

    nrows = 2 + rowof(elast) - rc0
    ready, released, held = [0] * S, [0] * 4, {}
    # the tasks of each Maxwell wave, in its order, and the next one's index;
    # a task waits until the slots of all its rows are released, writes them
    # (the fluid rows go into the ring as soon as they are loaded), then
    # publishes its rows one at a time
    ptask = [[[k for k in range(t, t + task) if k < nrows] for t in range(task * m, nrows, task * nmax)]
             for m in range(nmax)]
    pk = [0] * nmax
    pstate = [0] * nmax             # 0: before the writes, j >= 1: written, j - 1 rows published
    ce = [z0 - 3] * 4               # next row e of each sweep wave
    cstate = [0] * 4                # 0: before the read, 1: read, not yet released
    reading = [None] * 4            # ring rows a sweep wave reads (between read and release)
    for _ in range(10 ** 6):
        actors = [("P", m) for m in range(nmax) if pk[m] < len(ptask[m])] + \
                 [("C", w) for w in range(4) if ce[w] <= elast]
        if not actors:
            return
        kind, i = rnd.choice(actors)
        if kind == "P":
            rows = ptask[i][pk[i]]
            if pstate[i] == 0:
                if rows[-1] >= S and min(released) < rows[-1] - S + 1:
                    continue                               # waits
                for k in rows:
                    old = held.get(k % S)
                    for w in range(4):
                        if old is not None and reading[w] is not None:
                            assert old not in reading[w], f"row {old} overwritten while wave {w} reads it"
                        if old is not None and ce[w] <= elast:
                            kc = 1 + rowof(ce[w]) - rc0
                            assert old not in (kc, kc - 1), f"row {old} overwritten before wave {w} reads it"
                    held[k % S] = k
                pstate[i] = 1
            else:
                k = rows[pstate[i] - 1]
                ready[k % S] = k + 1
                pstate[i] += 1
                if pstate[i] > len(rows):
                    pstate[i] = 0
                    pk[i] += 1
        else:
            e = ce[i]
            rc = rowof(e)
            kc = 1 + rc - rc0
            if cstate[i] == 0:
                if ready[kc % S] < kc + 1 or ready[(kc - 1) % S] < kc:
                    continue                               # waits
                assert held.get(kc % S) == kc and held.get((kc - 1) % S) == kc - 1
                reading[i] = (kc, kc - 1)
                cstate[i] = 1
            else:
                released[i] = kc + (rowof(e + 1) != rc) - 1
                reading[i] = None
                cstate[i] = 0
                ce[i] += 1
    raise AssertionError("no progress")


@pytest.mark.parametrize("nz", [4, 5, 7, 16, 33, 100])
@pytest.mark.parametrize("nmax", [1, 2, 4])
@pytest.mark.parametrize("task", [1, 2])
def test_progress_counters_end_and_protect_rows(nz, nmax, task):
    # fixed chunk heights and heights sized at run time (any height >= 16)
    for tz in (16, 32, 64, 37, 147):
        for z0 in range(0, nz, tz):
            for seed in range(20):
                run(z0, tz, nz, nmax, seed, task)
