"""The float32 arithmetic behind the gather's row set-up (csrc/knn.hip
cell_bounds and small_divmod), emulated in numpy with the hardware
reciprocal taken at the correctly rounded value and one ulp either side:

* cell_bounds never excludes a position's cell as the grid hashes it
  (cell_coord: the correctly rounded quotient, truncated, clamped to
  [0, g - 1]), over random positions and every cell boundary with its one-ulp
  neighbours, so pruning with it cannot drop a photon;
* small_divmod equals integer division for every s < 2^20.
"""
import numpy as np

F = np.float32


def cell_coord(p, mn, cell, g):
    f = ((p - mn).astype(F) / cell).astype(F)
    inside = np.isfinite(f) & (f >= 1) & (f < g)
    tr = np.trunc(np.where(inside, f, 0)).astype(np.int64)
    return np.where(~(f < F(g)), g - 1, np.where(f < 1, 0, np.minimum(tr, g - 1)))


def cell_bounds(lo, hi, mn, rcell, g):
    tl = ((lo - mn).astype(F) * rcell).astype(F)
    th = ((hi - mn).astype(F) * rcell).astype(F)
    tl = (tl - ((np.abs(tl) * F(2.0 ** -20)).astype(F) + F(2.0 ** -20))).astype(F)
    th = (th + ((np.abs(th) * F(2.0 ** -20)).astype(F) + F(2.0 ** -20))).astype(F)
    top = F(g - 1)
    xl = np.floor(np.minimum(np.maximum(tl, F(0)), top)).astype(np.int64)
    xh = np.floor(np.maximum(np.minimum(th, top), F(0))).astype(np.int64)
    return xl, xh


def test_cell_bounds_contain_the_hashed_cell():
    rng = np.random.default_rng(7)
    for _ in range(20):
        mn = F(rng.uniform(-5, 5))
        cell = F(10 ** rng.uniform(-4, 0))
        g = int(rng.integers(1, 3000))
        p = (mn + rng.uniform(-0.2, 1.2, 50_000) * g * cell).astype(F)
        edges = (mn + np.arange(-2, g + 3).astype(F) * cell).astype(F)
        p = np.concatenate([p, edges, np.nextafter(edges, F(-np.inf)), np.nextafter(edges, F(np.inf))]).astype(F)
        c = cell_coord(p, mn, cell, g)
        r0 = F(1) / cell
        for rcell in (r0, np.nextafter(r0, F(0)), np.nextafter(r0, F(np.inf))):
            xl, xh = cell_bounds(p, p, mn, rcell, g)
            assert np.all(xl <= c) and np.all(c <= xh)
            # and never more than one cell wider on either side
            assert np.all(c - xl <= 1) and np.all(xh - c <= 1)


def test_small_divmod_is_exact_below_2_20():
    s = np.arange(0, 1 << 20, dtype=np.int64)
    sf = (s.astype(F) + F(0.5)).astype(F)
    for d in list(range(1, 40)) + [63, 64, 65, 127, 255, 1000, 4095, 8193, 65535, 1 << 19]:
        r0 = F(1) / F(d)
        for rd in (r0, np.nextafter(r0, F(0)), np.nextafter(r0, F(2))):
            q = (sf * rd).astype(F).astype(np.int64)
            assert np.array_equal(q, s // d), d
            assert np.array_equal(s - q * d, s % d), d
