"""Pins for the oracle (CPU): known answers computed independently (decimal
arithmetic), closed forms, and physical invariants chosen so that a dropped
term, a wrong sign or an off-by-one slab index fails one of them."""
import numpy as np
import pytest

from oracle import fdtd_oracle as O


# --- scalars and 1-D profiles ------------------------------------------------
def test_cfl_dt_known_answer():
    # dx / (c sqrt 2) for dx = 0.05 m, computed with 40-digit decimals.
    assert O.cfl_dt(0.05) == pytest.approx(1.179327168374842045560806539877746e-10, rel=1e-15)


def test_n_iterations_matches_notebook_run():
    # example.ipynb: time_windows 1e-7, dt 1e-10 -> 1001 steps.
    assert O.n_iterations(1e-7, 1e-10) == 1001


def test_resolve_dt_clamps_to_cfl():
    lim = O.cfl_dt(0.05)
    assert O.resolve_dt(0.05, None) == lim
    assert O.resolve_dt(0.05, 2 * lim) == lim
    assert O.resolve_dt(0.05, 1e-10) == 1e-10


def test_ricker_peak_symmetry_and_t0_value():
    f = 1e8
    chi = np.sqrt(2.0) / f
    w = O.ricker(201, chi / 100, f)
    assert w[100] == pytest.approx(1.0, abs=1e-15)           # w(chi) = amp
    np.testing.assert_allclose(w[:100], w[200:100:-1], rtol=1e-12, atol=1e-15)
    # zero crossings at tau = +-1/(pi f sqrt 2)
    tz = chi - 1.0 / (np.pi * f * np.sqrt(2.0))
    wz = O.ricker(2, tz, f)[1]
    assert abs(wz) < 1e-12
    # w(0) = (1 - 4 pi^2) exp(-2 pi^2), decimal known answer
    assert O.ricker(1, 1e-10, f)[0] == pytest.approx(-1.029408485324770591810513904e-7, rel=1e-12)
    assert O.ricker(1, 1e-10, f, dtype=np.float32)[0] == pytest.approx(-1.0294e-7, rel=2e-3)
    w2 = O.ricker(50, 1e-10, f, amp=2.5)
    np.testing.assert_allclose(w2, 2.5 * O.ricker(50, 1e-10, f), rtol=1e-15)


def test_cpml_profile_known_answers_P2():
    pro = O.cpml_profile(2, 0.05, 1e-10, np.full(7, 4.0), np.ones(7))
    assert pro["smax"] == pytest.approx(0.1061767491916050660283435881, rel=1e-13)
    kat = {  # (RA, RE, RF) per q, from 40-digit decimal arithmetic
        "E": [(1.0, 1.0, 0.0),
              (0.9638795235523192201865531709, 0.9277590471046384403731063419,
               0.07224095289536155962689365810)],
        "H": [(0.9976633441702724416407903531, 0.9953266883405448832815807063,
               0.004673311659455116718419293734),
              (0.8405392660751235756951642712, 0.6810785321502471513903285424,
               0.3189214678497528486096714576)],
    }
    for tag, rows in kat.items():
        for q, (ra, re, rf) in enumerate(rows):
            assert pro[tag + "RA"][q] == pytest.approx(ra, rel=1e-13, abs=1e-16)
            assert pro[tag + "RB"][q] == pytest.approx(ra, rel=1e-13, abs=1e-16)
            assert pro[tag + "RE"][q] == pytest.approx(re, rel=1e-13, abs=1e-16)
            assert pro[tag + "RF"][q] == pytest.approx(rf, rel=1e-13, abs=1e-16)


def test_cpml_identities_random():
    rng = np.random.default_rng(3)
    for P in (1, 5, 20):
        pro = O.cpml_profile(P, rng.uniform(0.005, 0.1), rng.uniform(1e-11, 1e-10),
                             rng.uniform(1, 20, 9), rng.uniform(1, 2, 9))
        for t in "EH":
            ra, rb, re, rf = (pro[t + k] for k in ("RA", "RB", "RE", "RF"))
            np.testing.assert_array_equal(ra, rb)
            np.testing.assert_allclose(re + rf, 1.0, atol=1e-14)
            np.testing.assert_allclose(ra + rf / 2, 1.0, atol=1e-14)
        assert pro["ERF"][0] == 0.0 and np.all(np.diff(pro["ERF"]) > 0)
        assert np.all(pro["HRF"] > pro["ERF"]) if P > 1 else True


def test_cpml_sigma_max_uses_first_row_mean():
    a = O.cpml_profile(4, 0.05, 1e-10, np.array([2.0, 6.0]), np.array([1.0, 1.0]))
    b = O.cpml_profile(4, 0.05, 1e-10, np.array([4.0, 4.0]), np.array([1.0, 1.0]))
    assert a["smax"] == b["smax"]


# --- per-cell coefficients ---------------------------------------------------
def test_cell_coeffs_known_answers():
    er, se, mr = np.full((2, 3), 4.0), np.full((2, 3), 0.01), np.full((2, 3), 2.0)
    c = O.cell_coeffs(er, se, mr, 1e-10, 0.05)
    kat = {"uE0": 0.9721578381907080361053298315, "uE1": 55.68432361858392778934033703,
           "uE4": 2.784216180929196389467016851, "uH1": 0.0007957747150262762457000069369,
           "uH4": 0.00003978873575131381228500034684, "uH0": 1.0}
    for k, v in kat.items():
        a = getattr(c, k)
        assert a.shape == (3, 4)
        np.testing.assert_allclose(a[:2, :3], v, rtol=1e-13)
        assert np.all(a[2, :] == 0) and np.all(a[:, 3] == 0)   # padding (uc.cu:19)
    ca = O.cell_coeffs(er, se, mr, 1e-10, 0.05, adjoint=True)
    np.testing.assert_allclose(ca.uE0[:2, :3], 0.9860789190953540180526649157, rtol=1e-13)
    np.testing.assert_array_equal(ca.uE1, c.uE1)
    np.testing.assert_array_equal(ca.uH4, c.uH4)


def test_cell_coeffs_pec_switch_and_float32_mode():
    se = np.array([[0.0, 100.0, 100.5]])
    c = O.cell_coeffs(np.ones((1, 3)), se, np.ones((1, 3)), 1e-10, 0.05)
    assert c.uE0[0, 0] == 1.0 and c.uE0[0, 1] != 0       # 100 is not > 100
    assert c.uE0[0, 2] == 0 and c.uE1[0, 2] == 0 and c.uE4[0, 2] == 0
    assert c.uH1[0, 2] > 0
    c32 = O.cell_coeffs(np.full((1, 1), 6.0), np.full((1, 1), 2e-3), np.ones((1, 1)),
                        1e-10, 0.05, dtype=np.float32)
    c64 = O.cell_coeffs(np.full((1, 1), 6.0), np.full((1, 1), 2e-3), np.ones((1, 1)),
                        1e-10, 0.05)
    assert c32.uE4.dtype == np.float32
    np.testing.assert_allclose(c32.uE4[0, 0], c64.uE4[0, 0], rtol=1e-6)


# --- forward invariants --------------------------------------------------------
def _box(NI=40, NJ=34, P=0, er=None, se=None, mr=None):
    er = np.full((NI, NJ), 4.0) if er is None else er
    se = np.zeros((NI, NJ)) if se is None else se
    mr = np.ones((NI, NJ)) if mr is None else mr
    return er, se, mr


def _pulse(nt, n_on=30):
    s = np.zeros((1, nt))
    s[0, :n_on] = np.sin(np.pi * np.arange(n_on) / n_on) ** 2
    return s


def test_lossless_box_conserves_discrete_energy():
    """Yee leapfrog: Q^n = sum Ez^2/uE1 + sum H^{n-1/2} H^{n+1/2}/uH1 is exactly
    conserved once the source stops (heterogeneous eps and mu)."""
    rng = np.random.default_rng(1)
    NI, NJ = 36, 30
    er = rng.uniform(1, 9, (NI, NJ))
    mr = rng.uniform(1, 2, (NI, NJ))
    c = O.cell_coeffs(er, np.zeros((NI, NJ)), mr, 0.9 * O.cfl_dt(0.05), 0.05)
    nt = 1500
    Q = []
    prev = {}

    def hook(n, st):
        if "Hx" in prev:
            live_e = c.uE1 > 0
            live_h = c.uH1 > 0
            q = np.sum(prev["Ez"][0][live_e] ** 2 / c.uE1[live_e])
            q += np.sum(prev["Hx"][0][live_h] * st.Hx[0][live_h] / c.uH1[live_h])
            q += np.sum(prev["Hy"][0][live_h] * st.Hy[0][live_h] / c.uH1[live_h])
            Q.append(q)
        prev["Hx"], prev["Hy"], prev["Ez"] = st.Hx.copy(), st.Hy.copy(), st.Ez.copy()

    O.forward(c, {}, 0, 0.05, np.array([[[18, 15]]]), _pulse(nt) * 1e3, np.array([[[5, 5]]]),
              nt, store=False, hook=hook)
    Q = np.array(Q[40:])
    assert Q[0] > 0
    np.testing.assert_allclose(Q, Q[0], rtol=1e-10)


def test_lossy_box_energy_decays():
    er, se, mr = _box(se=np.full((40, 34), 5e-3))
    c = O.cell_coeffs(er, se, mr, 0.9 * O.cfl_dt(0.05), 0.05)
    peaks = []
    O.forward(c, {}, 0, 0.05, np.array([[[20, 17]]]), _pulse(1200) * 1e3, np.array([[[5, 5]]]),
              1200, store=False,
              hook=lambda n, st: peaks.append(np.sum(st.Ez ** 2)))
    late = np.array(peaks[600:]).reshape(-1, 100).max(axis=1)
    assert np.all(np.diff(late) < 0)


def _run(er, se, mr, P, src, rx, nt, dx=0.05, dtype=np.float64, wavelet=None):
    dt = 0.9 * O.cfl_dt(dx)
    w = O.ricker(nt, dt, 3e8) if wavelet is None else wavelet
    pb = O.Problem(er, se, mr, dx, dt, nt, P, np.asarray(src), np.asarray(rx), w[None, :])
    return O.run_forward(pb, dtype, store=False)[0]


@pytest.mark.parametrize("axis", [0, 1])
def test_mirror_symmetry_with_cpml(axis):
    """A model symmetric about the grid centre gives mirror-equal traces.  The
    E cells i <-> NI - i and H cells i <-> NI - 1 - i map the minus CPML slab
    onto the plus slab, so any slab index shift breaks this."""
    NI, NJ, P = 48, 44, 8
    rng = np.random.default_rng(2)
    if axis == 1:
        prof = rng.uniform(2, 8, NI)
        er = np.repeat(prof[:, None], NJ, axis=1)
        src = [[[17, NJ // 2]]]
        rx = [[[30, NJ // 2 - 9], [30, NJ // 2 + 9], [11, NJ // 2 - 15], [11, NJ // 2 + 15]]]
    else:
        prof = rng.uniform(2, 8, NJ)
        er = np.repeat(prof[None, :], NI, axis=0)
        src = [[[NI // 2, 15]]]
        rx = [[[NI // 2 - 9, 30], [NI // 2 + 9, 30], [NI // 2 - 17, 7], [NI // 2 + 17, 7]]]
    se = np.full((NI, NJ), 2e-3)
    tr = _run(er, se, np.ones((NI, NJ)), P, src, rx, 400)
    assert np.abs(tr).max() > 0
    np.testing.assert_allclose(tr[0, :, 0], tr[0, :, 1], rtol=1e-9, atol=1e-12 * np.abs(tr).max())
    np.testing.assert_allclose(tr[0, :, 2], tr[0, :, 3], rtol=1e-9, atol=1e-12 * np.abs(tr).max())


def test_source_receiver_reciprocity():
    NI, NJ = 40, 36
    er, se, mr = _box(NI, NJ, se=np.full((NI, NJ), 3e-3))
    a, b = (9, 11), (27, 24)
    tab = _run(er, se, mr, 0, [[a]], [[b]], 300)
    tba = _run(er, se, mr, 0, [[b]], [[a]], 300)
    np.testing.assert_allclose(tab, tba, rtol=1e-10, atol=1e-13 * np.abs(tab).max())


def test_linearity_in_source_amplitude():
    NI, NJ = 30, 28
    er, se, mr = _box(NI, NJ, P=5, se=np.full((NI, NJ), 1e-3))
    w = O.ricker(200, 0.9 * O.cfl_dt(0.05), 3e8)
    for dt_ in (np.float64, np.float32):
        t1 = _run(er, se, mr, 5, [[(14, 13)]], [[(20, 20)]], 200, dtype=dt_, wavelet=w)
        t2 = _run(er, se, mr, 5, [[(14, 13)]], [[(20, 20)]], 200, dtype=dt_, wavelet=2 * w)
        np.testing.assert_array_equal(t2, 2 * t1)


def test_cpml_absorbs_against_larger_domain():
    """Traces next to a 20-cell CPML agree with a 4x larger domain (no
    boundary within reach during the window) to better than -40 dB."""
    n, P, big = 80, 20, 320
    nt = 380
    er_s = np.full((n, n), 4.0)
    se_s = np.full((n, n), 1e-3)
    off = (big - n) // 2
    src, rx = (40, 40), [(40, 24), (26, 40), (52, 50)]
    small = _run(er_s, se_s, np.ones((n, n)), P, [[src]], [rx], nt)
    er_b = np.full((big, big), 4.0)
    se_b = np.full((big, big), 1e-3)
    shift = lambda p: (p[0] + off, p[1] + off)
    large = _run(er_b, se_b, np.ones((big, big)), P, [[shift(src)]], [[shift(p) for p in rx]], nt)
    err = np.linalg.norm(small - large) / np.linalg.norm(large)
    assert err < 1e-2, err
    # without CPML the same comparison fails: the test can tell
    small_pec = _run(er_s, se_s, np.ones((n, n)), 0, [[src]], [rx], nt)
    assert np.linalg.norm(small_pec - large) / np.linalg.norm(large) > 0.1


def test_fp32_mode_tracks_fp64():
    NI, NJ = 50, 40
    rng = np.random.default_rng(4)
    er = rng.uniform(3, 9, (NI, NJ))
    se = rng.uniform(0, 5e-3, (NI, NJ))
    t64 = _run(er, se, np.ones((NI, NJ)), 8, [[(25, 20)]], [[(10, 30), (40, 10)]], 500)
    t32 = _run(er, se, np.ones((NI, NJ)), 8, [[(25, 20)]], [[(10, 30), (40, 10)]], 500,
               dtype=np.float32)
    err = np.linalg.norm(t32 - t64) / np.linalg.norm(t64)
    assert 0 < err < 1e-5


# --- reference-mode adjoint invariants -------------------------------------------
def _adj_case(nt=160):
    NI, NJ, P = 36, 32, 6
    rng = np.random.default_rng(9)
    er = rng.uniform(3, 8, (NI, NJ))
    se = rng.uniform(0, 4e-3, (NI, NJ))
    dt = 0.9 * O.cfl_dt(0.05)
    pb = O.Problem(er, se, np.ones((NI, NJ)), 0.05, dt, nt, P,
                   np.array([[[15, 10]], [[20, 12]]]), np.array([[[25, 20], [8, 22]]] * 2),
                   O.ricker(nt, dt, 3e8)[None, :])
    tr, snaps = O.run_forward(pb)
    resid = rng.standard_normal(tr.shape)
    return pb, snaps, resid


def test_adjoint_reference_linear_in_residual_and_normalised():
    pb, snaps, resid = _adj_case()
    ge, gs, ge_raw, gs_raw = O.run_gradient(pb, resid, snaps)
    _, _, ge2, gs2 = O.run_gradient(pb, 2 * resid, snaps)
    np.testing.assert_array_equal(ge2, 2 * ge_raw)
    np.testing.assert_array_equal(gs2, 2 * gs_raw)
    assert np.linalg.norm(ge) == pytest.approx(1.0, rel=1e-12)
    assert np.linalg.norm(gs) == pytest.approx(1.0, rel=1e-12)
    # additivity over residuals
    r2 = np.roll(resid, 7, axis=1)
    _, _, ga, _ = O.run_gradient(pb, r2, snaps)
    _, _, gb, _ = O.run_gradient(pb, resid + r2, snaps)
    np.testing.assert_allclose(gb, ge_raw + ga, rtol=1e-9, atol=1e-12 * np.abs(gb).max())


def test_adjoint_reference_zero_on_dead_lines():
    pb, snaps, resid = _adj_case()
    _, _, ge, gs = O.run_gradient(pb, resid, snaps)
    assert np.abs(ge).max() > 0
    for g in (ge, gs):
        assert np.all(g[0, :] == 0) and np.all(g[:, 0] == 0)


# --- exact adjoint: dot-product test and central finite differences (FP64) ----
def _exact_case():
    rng = np.random.default_rng(0)
    NI, NJ, P = 30, 26, 5
    er = rng.uniform(2, 8, (NI, NJ))
    se = rng.uniform(1e-3, 5e-3, (NI, NJ))
    mr = np.ones((NI, NJ))
    dt = 0.9 * O.cfl_dt(0.05)
    nt = 150
    src = np.array([[[12, 9]], [[17, 15]]])
    rx = np.array([[[20, 18], [6, 20], [6, 20]]] * 2)       # includes a colliding pair
    w = O.ricker(nt, dt, 4e8)[None, :]

    def F(er_, se_):
        return O.run_forward(O.Problem(er_, se_, mr, 0.05, dt, nt, P, src, rx, w))

    tr, sn = F(er, se)
    r = rng.standard_normal(tr.shape)
    ge, gs = O.run_gradient_exact(O.Problem(er, se, mr, 0.05, dt, nt, P, src, rx, w), r, sn)
    return rng, er, se, F, r, ge, gs


def test_exact_adjoint_dot_product_with_finite_differences():
    """<J dm, r> by central differences of the forward map equals <dm, J^T r>
    (row 0 is held fixed: it also enters the CPML sigma_max)."""
    rng, er, se, F, r, ge, gs = _exact_case()
    for name, g, base in (("er", ge, er), ("se", gs, se)):
        d = rng.standard_normal(base.shape)
        d[0, :] = 0
        h = 1e-4 * np.abs(base).mean()
        if name == "er":
            fp, fm = F(er + h * d, se)[0], F(er - h * d, se)[0]
        else:
            fp, fm = F(er, se + h * d)[0], F(er, se - h * d)[0]
        fd = np.sum((fp - fm) / (2 * h) * r)
        ad = np.sum(d * g)
        assert abs(fd - ad) <= 1e-5 * abs(fd), (name, fd, ad)


def test_exact_adjoint_single_cell_finite_differences():
    rng, er, se, F, r, ge, gs = _exact_case()
    cells = [(int(rng.integers(1, 30)), int(rng.integers(0, 26))) for _ in range(6)]
    for (i, j) in cells:
        for name, g, base in (("er", ge, er), ("se", gs, se)):
            h = 1e-3 * base[i, j]
            p = base.copy(); p[i, j] += h
            m = base.copy(); m[i, j] -= h
            if name == "er":
                fd = np.sum((F(p, se)[0] - F(m, se)[0]) / (2 * h) * r)
            else:
                fd = np.sum((F(er, p)[0] - F(er, m)[0]) / (2 * h) * r)
            assert abs(fd - g[i, j]) <= 1e-5 * max(abs(fd), 1e-12 * np.abs(g).max()), (name, i, j)
