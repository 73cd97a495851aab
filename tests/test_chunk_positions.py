"""The finishing pass's chunk positions (kinectfusion-cuda_amd/csrc/raycast.hip,
chunk_position) against the chained additions they replace.

The march takes its sample positions by repeated rounded addition, t_{k+1} =
fl(t_k + step) (raycast.cu:75-145 adds the step to t once per sample).  The
finishing pass needs t_0 .. t_63 of a chunk at once, one per lane.  Each lane
guesses t_{k-1} on the line t_b + (k-1-b) d through an exact t_b, takes one
rounded step from the guess, and checks that the result lies on the line too;
the first lane whose check fails is exact and becomes the next t_b.  This test
runs that scheme (numpy float32 arithmetic; fma emulated exactly in float64)
against the sequential chain over random and adversarial (t, step) pairs:
binade crossings, rounding ties, steps below half an ulp, large steps."""
import numpy as np

LANES = 64


def fma32(a, b, c):
    # exact for these operands: |a| <= 64 and b, c float32 of similar scale keep the
    # float64 sum exact (well below 53 significant bits), then one rounding to float32
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def sequential(t, step):
    out = np.empty((t.size, LANES), np.float32)
    cur = t.copy()
    for k in range(LANES):
        out[:, k] = cur
        cur = (cur + step).astype(np.float32)
    return out


def chunk_positions(t, step):
    """The device scheme, vectorised over independent chunks (one row per chunk)."""
    n = t.size
    lane = np.arange(LANES, dtype=np.int64)[None, :]
    res = np.repeat(t[:, None], LANES, axis=1).astype(np.float32)
    b = np.zeros(n, np.int64)
    tb = t.copy()
    active = np.ones(n, bool)
    segments = np.zeros(n, np.int64)
    while active.any():
        idx = np.flatnonzero(active)
        segments[idx] += 1
        s = step[idx][:, None]
        tbb = tb[idx][:, None]
        bb = b[idx][:, None]
        d = ((tbb + s).astype(np.float32) - tbb).astype(np.float32)
        c = fma32((lane - bb - 1).astype(np.float32), d, tbb)
        tk = (c + s).astype(np.float32)
        on = tk == fma32((lane - bb).astype(np.float32), d, tbb)
        upd = lane > bb
        r = res[idx]
        r[upd] = tk[upd]
        res[idx] = r
        off = upd & ~on
        has = off.any(axis=1)
        first = np.argmax(off, axis=1)
        done = idx[~has]
        active[done] = False
        go = idx[has]
        b[go] = first[has]
        tb[go] = res[go, first[has]]
    return res, segments


def check(t, step):
    t = np.asarray(t, np.float32)
    step = np.asarray(step, np.float32)
    ref = sequential(t, step)
    got, seg = chunk_positions(t, step)
    bad = np.flatnonzero((ref.view(np.uint32) != got.view(np.uint32)).any(axis=1))
    assert bad.size == 0, (t[bad[:5]], step[bad[:5]])
    return seg


def test_random_march_ranges():
    rng = np.random.default_rng(0)
    n = 200000
    t = rng.uniform(0.3, 8.0, n).astype(np.float32)
    mue = rng.uniform(0.005, 0.1, n).astype(np.float32)
    small = (mue * np.float32(0.1)).astype(np.float32)
    large = (mue * np.float32(1.99)).astype(np.float32)
    seg_small = check(t, small)
    seg_large = check(t, large)
    # a chunk of small steps is nearly always one binade: few segments
    assert np.mean(seg_small) < 3.0 and np.mean(seg_large) < 8.0, (np.mean(seg_small), np.mean(seg_large))


def test_binade_crossings():
    rng = np.random.default_rng(1)
    e = rng.integers(-2, 4, 100000)
    s = rng.uniform(1e-4, 0.2, e.size).astype(np.float32)
    k = rng.integers(0, 64, e.size)
    t = (np.ldexp(1.0, e) - k * s.astype(np.float64)).astype(np.float32)
    t = np.maximum(t, np.float32(1e-3))
    check(t, s)


def test_rounding_ties():
    # step = (m + 1/2) ulp(t): every addition is a tie until the parity settles
    rng = np.random.default_rng(2)
    e = rng.integers(-2, 3, 100000)
    ulp = np.ldexp(1.0, e - 23)
    m = rng.integers(1, 1 << 18, e.size)
    s = ((m + 0.5) * ulp).astype(np.float32)
    assert np.all(s.astype(np.float64) == (m + 0.5) * ulp)
    t = (np.ldexp(1.0, e) * rng.uniform(1.0, 1.9, e.size)).astype(np.float32)
    check(t, s)


def test_steps_below_half_an_ulp_and_large_steps():
    rng = np.random.default_rng(3)
    t = rng.uniform(0.5, 6.0, 50000).astype(np.float32)
    tiny = (np.spacing(t) * rng.uniform(0.01, 0.5, t.size)).astype(np.float32)  # t + s rounds back to t
    check(t, tiny)
    big = (t * rng.uniform(0.1, 3.0, t.size)).astype(np.float32)
    check(t, big)
