"""Model check of the progress-counter protocol that couples the five waves
of k_mxz (Maxwell + the next step's sweeps in one launch): the Maxwell wave
writes ring row k into slot k % S after every sweep wave has released row
k - S, and publishes `produced = k + 1`; a sweep wave at ring row kc waits for
produced >= kc + 1, reads rows kc and kc - 1, then releases the rows below the
ones its next row reads.  Random interleavings of the five waves over the
row sequences of many chunk shapes must finish (no wait without end) and
never overwrite a row that a sweep wave may still read."""
import random

import pytest

S = 6       # MXZ_SLOTS


def run(z0, tz, nz, seed):
    rnd = random.Random(seed)
    zend = min(z0 + tz, nz)
    elast = zend + 2
    rc0 = max(0, min(z0 - 3, nz - 1))

    def rowof(e):
        return max(0, min(min(e, elast), nz - 1))

    nrows = 2 + rowof(elast) - rc0
    produced, released, held = 0, [0] * 4, {}
    pk, ce = 0, [z0 - 3] * 4
    for _ in range(10 ** 6):
        actors = (["P"] if pk < nrows else []) + [w for w in range(4) if ce[w] <= elast]
        if not actors:
            return
        a = rnd.choice(actors)
        if a == "P":
            k = pk
            if k >= S and min(released) < k - S + 1:
                continue                                   # waits
            old = held.get(k % S)
            for w in range(4):
                if ce[w] <= elast and old is not None:
                    kc = 1 + rowof(ce[w]) - rc0
                    assert old not in (kc, kc - 1), f"row {old} overwritten while wave {w} reads it"
            held[k % S] = k
            produced, pk = k + 1, pk + 1
        else:
            e = ce[a]
            rc = rowof(e)
            kc = 1 + rc - rc0
            if produced < kc + 1:
                continue                                   # waits
            assert held.get(kc % S) == kc and held.get((kc - 1) % S) == kc - 1
            released[a] = kc + (rowof(e + 1) != rc) - 1
            ce[a] += 1
    raise AssertionError("no progress")


@pytest.mark.parametrize("nz", [4, 5, 7, 16, 33, 100])
def test_progress_counters_end_and_protect_rows(nz):
    for tz in (16, 32, 64):
        for z0 in range(0, nz, tz):
            for seed in range(20):
                run(z0, tz, nz, seed)
