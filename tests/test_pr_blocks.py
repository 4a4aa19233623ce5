"""The propagation layout of the compensated PageRank on a PE block, checked on
the host.

dmm_pr_blocks_check builds the layout the GPU kernels read (source windows of
32768 contributions, row buckets), checks every index the kernels use against
its buffer, and replays one iteration's per-row sums with the kernels'
arithmetic: each contribution as an integer multiple of 2^-80, summed exactly,
the sum rounded once to FP32.  Every row's result must therefore equal the
FP32 rounding (to nearest, ties to even) of the exact sum of its in-neighbours'
contributions, computed here with Python integers.  Graphs: a Kronecker PE
block, many source windows, rows without in-edges, a hub row with more
in-edges than one bucket holds (summed in parts), negative and zero
contributions, and an empty block."""
import ctypes as C

import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L


def f32_round(x):
    """FP32 rounding to nearest (ties to even) of the integer x * 2^-149."""
    neg = x < 0
    x = -x if neg else x
    if x == 0:
        return np.float32(0.0)
    nb = x.bit_length()
    if nb > 24:
        shift = nb - 24
        q, r = divmod(x, 1 << shift)
        half = 1 << (shift - 1)
        if r > half or (r == half and q & 1):
            q += 1
        v = np.float32(float(q) * 2.0 ** (shift - 149))
    else:
        v = np.float32(float(x) * 2.0 ** -149)
    return -v if neg else v


def exact_rows(oo, on, contrib):
    """Per row: the FP32 rounding of the exact sum of contrib[on[e]]."""
    units = [int(v) for v in (contrib.astype(np.float64) * 2.0 ** 149).tolist()]  # exact: power-of-two scaling
    out = np.empty(len(oo) - 1, np.float32)
    for r in range(len(oo) - 1):
        out[r] = f32_round(sum(units[u] for u in on[oo[r]:oo[r + 1]].tolist()))
    return out


def check(oo, on, n_gather, contrib):
    lib = L.load()
    n = len(oo) - 1
    oo = np.ascontiguousarray(oo, np.int64)
    on = np.ascontiguousarray(on, np.int32)
    contrib = np.ascontiguousarray(contrib, np.float32)
    inc = np.full(max(1, n), np.nan, np.float32)
    stats = np.zeros(6, np.int64)
    rc = lib.dmm_pr_blocks_check(n, n_gather, L.ptr(oo, C.c_int64), L.ptr(on, C.c_int32),
                                 L.ptr(contrib, C.c_float), L.ptr(inc, C.c_float), L.ptr(stats, C.c_int64))
    if rc:
        raise AssertionError(lib.dmm_last_error(None).decode())
    assert np.array_equal(inc[:n], exact_rows(oo, on, contrib))
    assert stats[0] == max(1, -(-n_gather // 32768))
    return stats


def _contrib(rng, k, lo=1e-9, hi=1e-3):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=k)).astype(np.float32)


def test_blocks_kronecker_pe_block(dmm, orc):
    g = dmm.build_graph(13, 16, uniform=False)
    og = orc.CSR.of(g)
    p = dmm.partition(og.n, 8, 3)
    b, e = p["start"], p["end"]
    oo = og.io[b:e + 1] - og.io[b]
    on = og.inn[og.io[b]:og.io[e]]
    st = check(oo, on, og.n, _contrib(np.random.default_rng(1), og.n))
    assert st[2] >= 1 and st[3] == 0


def test_blocks_many_windows_and_empty_rows():
    """Sources over seven windows (the last one short), a third of the rows
    without in-edges, negative and zero contributions."""
    rng = np.random.default_rng(2)
    n, n_gather = 20000, 200001
    deg = rng.integers(0, 40, size=n) * (rng.random(n) < 0.67)
    oo = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=oo[1:])
    on = rng.integers(0, n_gather, size=int(oo[-1])).astype(np.int32)
    on = np.concatenate([np.sort(on[oo[r]:oo[r + 1]]) for r in range(n)])
    contrib = _contrib(rng, n_gather)
    contrib[rng.random(n_gather) < 0.1] *= -1
    contrib[rng.random(n_gather) < 0.05] = 0
    st = check(oo, on, n_gather, contrib)
    assert st[0] == 7 and st[1] == 3 and st[3] == 0  # 20000 rows: buckets of 8192


def test_blocks_hub_row_summed_in_parts():
    """A row with more in-edges than one bucket holds (2^19) is summed in
    parts; the rows around it stay in ordinary buckets."""
    rng = np.random.default_rng(3)
    n_gather = 100000
    degs = [7, 0, 600000, 12, 3]
    oo = np.zeros(len(degs) + 1, np.int64)
    np.cumsum(degs, out=oo[1:])
    on = np.concatenate([np.sort(rng.integers(0, n_gather, size=d)) for d in degs]).astype(np.int32)
    st = check(oo, on, n_gather, _contrib(rng, n_gather))
    assert st[3] == 1 and st[5] == 2  # one row in two parts
    assert st[1] == 3  # rows {0, 1}, {2}, {3, 4}


def test_blocks_empty():
    check(np.zeros(5, np.int64), np.zeros(0, np.int32), 10, np.ones(10, np.float32))
