"""The gather walk's packed stream (csrc/knn.hip walk2), modelled step by
step on the CPU: the runs of a 64-row batch are laid end to end, a step takes
the next 128 stream positions (lane l: positions p + l and p + 64 + l), each
lane turns its position into a photon slot through the byte offset of the run
holding it (updated only for the runs that start inside the step, with the
same wrap-around unsigned arithmetic), and lanes past the stream are masked.
Every candidate of every run must come out exactly once and in run order,
which is what keeps the kernel's buffer contents, and so its results, those
of a run-by-run walk."""
import numpy as np

M = 2 ** 32


def packed_walk(b, ln):
    """Slots visited, in order, by the packed walk over runs (b[j], ln[j])."""
    incl = np.cumsum(ln).astype(np.int64)
    total = int(incl[-1])
    if total == 0:
        return []
    pre = incl - ln
    d4 = ((b - pre) * 4) % M
    rows = [j for j in range(len(ln)) if ln[j] > 0]
    state = {"nstart": 0, "cur": 0}
    lane = np.arange(64)

    def load(p):
        da = np.full(64, state["cur"], np.int64)
        db = da.copy()
        while rows and state["nstart"] < p + 128:
            j = rows.pop(0)
            dj = int(d4[j])
            t0 = state["nstart"] - p
            da = np.where(lane >= t0, dj, da)
            db = np.where(lane >= t0 - 64, dj, db)
            state["cur"] = dj
            state["nstart"] = int(pre[rows[0]]) if rows else 0xFFFFFFFF
        left = total - p
        oa = (da + lane * 4 + p * 4) % M
        ob = (db + lane * 4 + p * 4 + 256) % M
        return lane < left, lane + 64 < left, oa, ob

    out = []
    p = 0
    step = load(0)
    while True:
        pn = p + 128
        more = pn < total
        nxt = load(pn) if more else None
        va, vb, oa, ob = step
        out += [int(o) // 4 for o in oa[va]] + [int(o) // 4 for o in ob[vb]]
        p, step = pn, nxt
        if not more:
            return out


def test_packed_walk_visits_every_candidate_once_in_order():
    rng = np.random.default_rng(3)
    edge = [0, 1, 63, 64, 65, 127, 128, 129, 255, 256]
    for it in range(3000):
        nr = int(rng.integers(1, 65))
        kind = it % 4
        ln = np.zeros(64, np.int64)
        for j in range(nr):
            ln[j] = (rng.integers(0, 5), rng.integers(0, 200), rng.integers(0, 600), rng.choice(edge))[kind]
        b = rng.integers(0, 1_000_000, 64).astype(np.int64)
        want = [int(b[j]) + t for j in range(64) for t in range(int(ln[j]))]
        assert packed_walk(b, ln) == want, (it, kind)


def test_packed_walk_slot_below_stream_start():
    # a run whose first slot is below its stream start (b_j < pre_j) wraps
    # the offset around 2^32; the slots still come out right
    ln = np.array([300, 5, 0, 90] + [0] * 60, np.int64)
    b = np.array([1000, 3, 0, 7] + [0] * 60, np.int64)
    want = list(range(1000, 1300)) + list(range(3, 8)) + list(range(7, 97))
    assert packed_walk(b, ln) == want
