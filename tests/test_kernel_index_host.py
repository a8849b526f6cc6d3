"""The time-blocked kernels' integer arithmetic, checked on the host (no GPU)
through the library's host-compiled copies of the same functions.

* ``tb_item`` (the work order of a launch): over many random geometries
  (grid, CPML, shots, variant), every workgroup maps to a distinct
  (shot, tile) and every pair is covered -- a bijection -- and the items the
  launch runs first are exactly the interior tiles: regions that touch no
  CPML cell and no domain edge.
* ``FastDiv`` (multiply-and-shift division by a launch constant): exact
  against integer division, exhaustively for the divisors and dividends a
  launch can produce up to large grids, and at the ends of the 31-bit range.
"""
import ctypes

import numpy as np
import pytest

from fastgpr import _native
from fastgpr._native import Problem, lib

VARIANTS = [(32, 256, 1), (32, 256, 6), (64, 512, 8), (64, 256, 4), (32, 512, 6), (64, 1024, 6),
            (64, 1024, 8), (64, 1024, 2)]


def _problem(NI, NJ, P, S, tb):
    """A descriptor with non-null stand-in pointers (nothing is read or
    launched: the mapping only looks at sizes)."""
    p = Problem()
    p.NI, p.NJ, p.P, p.S, p.nt, p.nsrc, p.nrx = NI, NJ, P, S, 10, 0, 0
    p.pitch = _native.pitch(NJ)
    p.dx, p.dt = 0.01, 1e-11
    fake = 256
    for f in ("coef", "pml", "Ez", "Hx", "Hy", "psi", "Ez2", "Hx2", "Hy2", "psi2", "gshot", "traces",
              "snaps"):
        setattr(p, f, fake)
    p.tb_ri, p.tb_nt, p.tb_k = tb
    _, _, ti, tj = _native.tb_tiles(NI, NJ, *tb)
    p.tb_ti, p.tb_tj = ti, tj
    return p


def _interior(NI, NJ, P, tb, tile, tj_count):
    """The interior rule from the geometry: the region (tile centre plus
    halo) starts at row >= P + 1 and ends at row <= NI - P, likewise for
    columns."""
    ri, _, k = tb
    TI, TJ = ri - 2 * k - 1, 64 - 2 * k - 1
    ti, tj = np.divmod(tile, tj_count)
    gi0, gj0 = ti * TI - k, tj * TJ - k
    return (gi0 >= P + 1) & (gi0 + ri <= NI - P) & (gj0 >= P + 1) & (gj0 + 64 <= NJ - P)


def _map(p, adj):
    n = p.tb_ti * p.tb_tj * p.S
    shot = np.empty(n, np.int32)
    tile = np.empty(n, np.int32)
    inter = np.empty(n, np.int8)
    got = lib.fgpr_tb_map(ctypes.byref(p), adj, shot.ctypes.data, tile.ctypes.data,
                          inter.ctypes.data, n)
    assert got == n, lib.fgpr_last_error()
    return shot, tile, inter.astype(bool)


def test_work_item_mapping_is_a_bijection_with_interior_items_first():
    rng = np.random.default_rng(7)
    cases = [(400, 200, 20, 13, (64, 1024, 6)), (1200, 600, 20, 15, (64, 1024, 6)),
             (3, 3, 0, 1, (64, 1024, 6)), (60, 44, 8, 3, (32, 512, 4))]
    while len(cases) < 400:
        P = int(rng.integers(0, 30))
        NI = int(rng.integers(2 * P + 2, 2 * P + 900))
        NJ = int(rng.integers(2 * P + 2, 2 * P + 500))
        tb = VARIANTS[int(rng.integers(len(VARIANTS)))]
        cases.append((NI, NJ, P, int(rng.integers(1, 40)), tb))
    for NI, NJ, P, S, tb in cases:
        p = _problem(NI, NJ, P, S, tb)
        ntiles = p.tb_ti * p.tb_tj
        for adj in (0, 1):
            shot, tile, inter = _map(p, adj)
            assert shot.min() >= 0 and shot.max() < S and tile.min() >= 0 and tile.max() < ntiles
            key = shot.astype(np.int64) * ntiles + tile
            assert np.unique(key).size == S * ntiles, (NI, NJ, P, S, tb)       # bijection
            geo = _interior(NI, NJ, P, tb, tile, p.tb_tj)
            assert np.array_equal(inter, geo), (NI, NJ, P, S, tb)
            # per XCD (workgroup wg on XCD wg % 8, in order wg // 8) the
            # interior items come first
            for x in range(8):
                seq = inter[x::8]
                if seq.size:
                    first_general = np.argmin(seq) if not seq.all() else seq.size
                    assert not seq[first_general:].any(), (NI, NJ, P, S, tb, x)


def test_fastdiv_is_exact_over_the_launch_ranges():
    """Divisors up to 1024 exhaustively over dividends below 2^20 (shots,
    tile counts and workgroup indices of launches up to a million
    workgroups), every divisor up to 2^16 on the top of the 31-bit range, and
    random large divisors."""
    for d in range(1, 1025):
        assert lib.fgpr_fastdiv_mismatches(d, 0, 1 << 20) == 0, d
    top = (1 << 31)
    for d in list(range(1, 1 << 16, 97)) + [(1 << k) + o for k in range(1, 31) for o in (-1, 0, 1)]:
        if d < 1:
            continue
        assert lib.fgpr_fastdiv_mismatches(d, top - 4096, top) == 0, d
        lo, hi = min(max(0, d * 3 - 2048), top), min(top, d * 3 + 2048)
        assert lib.fgpr_fastdiv_mismatches(d, lo, hi) == 0, d       # around multiples of d
    rng = np.random.default_rng(3)
    for d in rng.integers(1, 1 << 30, size=200):
        n0 = int(rng.integers(0, top - 100_000))
        assert lib.fgpr_fastdiv_mismatches(int(d), n0, n0 + 100_000) == 0, (d, n0)
    assert lib.fgpr_fastdiv_mismatches(0, 0, 10) == -1


def test_planted_division_error_is_caught():
    """The check has teeth: a multiplier one below the exact one (computed
    here in Python, the same formula) fails for some dividend."""
    def fastdiv(n, d, bump=0):
        l = 0
        while (1 << l) < d:
            l += 1
        m = ((1 << 32) * ((1 << l) - d)) // d + 1 + bump
        t = (n * (m & 0xFFFFFFFF)) >> 32
        return (t + n) >> l
    n = np.arange(0, 1 << 20, dtype=np.int64)
    for d in (3, 7, 13, 51, 1000):
        assert all(fastdiv(int(v), d) == v // d for v in n[::997])
    bad = sum(fastdiv(int(v), 51, bump=-1) != v // 51 for v in n[:: 7])
    assert bad > 0


def _decode(e):
    u = int(np.uint32(np.int32(e)))
    return u >> 26, (u >> 20) & 63, u & 0xFFFFF


@pytest.mark.parametrize("tb", [(64, 1024, 6), (64, 1024, 8), (32, 512, 4), (64, 512, 3)])
def test_receiver_blocks_match_a_brute_force_region_scan(tb):
    """solver.tile_tables against a scan of every region: the forward block of
    a (shot, tile) holds exactly the live receivers in the tile centre, the
    adjoint block exactly those on updated cells anywhere in the region;
    entries give the cell relative to the region and the receiver index, in
    receiver order (the order receivers on one cell are applied in), and the
    block is packed (empty slots only at the end)."""
    from fastgpr.solver import tile_tables
    rng = np.random.default_rng(sum(tb))
    NI, NJ, S, nrx = 157, 131, 3, 40
    rx = np.stack([rng.integers(0, NI + 1, (S, nrx)), rng.integers(0, NJ + 1, (S, nrx))], -1)
    rx[0, 3] = rx[0, 1]                                  # two receivers on one cell
    rx[1, 5] = rx[1, 6] = rx[1, 2]                       # three
    rx[2, 7, 0] = -1                                     # dead
    rx[2, :4] = [(NI, 5), (7, NJ), (0, 0), (NI - 1, NJ - 1)]   # edges
    t = tile_tables(rx, NI, NJ, tb)
    RI, _, K = tb
    TI, TJ, ti, tj, cap = t["TI"], t["TJ"], t["ti"], t["tj"], t["cap"]
    assert cap == 64
    fwd = t["fwd"].reshape(S, ti * tj, cap)
    adj = t["adj"].reshape(S, ti * tj, cap)
    for s in range(S):
        for a in range(ti):
            for b in range(tj):
                gi0, gj0 = a * TI - K, b * TJ - K
                want_f, want_a = [], []
                for r in range(nrx):
                    i, j = int(rx[s, r, 0]), int(rx[s, r, 1])
                    if i < 0:
                        continue
                    if a * TI <= i < (a + 1) * TI and b * TJ <= j < (b + 1) * TJ:
                        want_f.append((i - gi0, j - gj0, r))
                    if i < NI and j < NJ and gi0 <= i < gi0 + RI and gj0 <= j < gj0 + 64:
                        want_a.append((i - gi0, j - gj0, r))
                for blk, want in ((fwd[s, a * tj + b], want_f), (adj[s, a * tj + b], want_a)):
                    n = len(want)
                    assert [_decode(e) for e in blk[:n]] == want, (s, a, b)
                    assert np.all(blk[n:] == -1), (s, a, b)


def test_receiver_blocks_grow_to_128_and_refuse_more():
    from fastgpr.solver import tile_tables
    rx = np.zeros((1, 100, 2), np.int64)
    rx[0, :, 0] = 30
    rx[0, :, 1] = np.arange(100) % 40 + 5                # 100 receivers in one region
    t = tile_tables(rx, 200, 200, (64, 1024, 6))
    assert t["cap"] == 128 and t["adj"].size == 128 * t["ti"] * t["tj"]
    rx = np.zeros((1, 129, 2), np.int64)
    rx[0, :, 0] = 30
    rx[0, :, 1] = 20                                     # 129 on one cell
    with pytest.raises(ValueError, match="at most 128"):
        tile_tables(rx, 200, 200, (64, 1024, 6))
