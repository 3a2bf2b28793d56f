"""GPU checks of the grouped launch over many problems (raysht_batch_*): a batch of
plans of different l_max, N_theta, m sets and column counts gives, bit for bit, what
each plan gives on its own, in every precision; the first problem is also checked
against the oracle, and the batch round trip recorded into a hipGraph replays
exactly."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import dev, empty, host, rel_max

pytestmark = pytest.mark.gpu

# (l_max, N_theta, m set, n_r, n_fields): uneven l_max, N_theta = 3(l_max+1)/2 where
# even, partial tiles in every direction, one plan over a subset of m
PROBLEMS = [
    (63, 96, "all", 5, 3),
    (31, 48, "all", 32, 1),
    (127, 192, "balanced", 3, 2),
    (11, 18, "all", 1, 1),
    (95, 144, "subset", 7, 2),
    (0, 2, "all", 2, 1),
]


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def m_values(lmax, kind):
    if kind == "balanced":
        return O.m_order_balanced(lmax)
    if kind == "subset":
        return np.array([lmax, 0, lmax // 2, 3, 1], dtype=np.int32)
    return np.arange(lmax + 1, dtype=np.int32)


def make(ra, precision):
    plans, inputs = [], []
    for i, (lmax, nt, kind, nr, nf) in enumerate(PROBLEMS):
        mv = m_values(lmax, kind)
        plans.append(ra.LegendrePlan(nt, lmax, mv, nr, nf, precision=precision))
        inputs.append(O.synthetic_spectral(lmax, mv, nr, nf, seed=100 + i))
    return plans, inputs


@pytest.mark.parametrize("precision", ["fp64", "bf16x6", "bf16x3", "int8x6", "int8x3"])
def test_batch_equals_single_plans(ra, precision):
    plans, inputs = make(ra, precision)
    batch = ra.LegendreBatch(plans)
    assert len(batch) == len(plans)
    s = [dev(x) for x in inputs]
    p_b = [empty(p.hybrid_size) for p in plans]
    batch.stp(s, p_b)
    s_b = [empty(p.spectral_size) for p in plans]
    batch.pts(p_b, s_b)
    for i, p in enumerate(plans):
        p1 = empty(p.hybrid_size)
        p.stp(s[i], p1)
        assert np.array_equal(host(p_b[i]), host(p1)), f"StP problem {i}"
        s1 = empty(p.spectral_size)
        p.pts(p1, s1)
        assert np.array_equal(host(s_b[i]), host(s1)), f"PtS problem {i}"
    # problem 0 against the oracle
    lmax, nt, kind, nr, nf = PROBLEMS[0]
    mv = m_values(lmax, kind)
    tabs = O.legendre_tables(nt, lmax, mv)
    p_ref = O.stp(nt, lmax, mv, nr, nf, tabs, inputs[0])
    tol = {"fp64": 1e-13, "bf16x6": 1e-12, "bf16x3": 1e-6, "int8x6": 1e-12, "int8x3": 1e-6}[precision]
    assert rel_max(host(p_b[0]), p_ref) <= tol


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_batch_schedules_and_autotune(ra, precision):
    """Every schedule candidate gives the same bits; the autotuner times the distinct
    candidates, keeps one of them, and the results stay the same."""
    plans, inputs = make(ra, precision)
    batch = ra.LegendreBatch(plans)
    s = [dev(x) for x in inputs]
    want_p, want_s = [], []
    for i, p in enumerate(plans):
        p1 = empty(p.hybrid_size)
        p.stp(s[i], p1)
        s1 = empty(p.spectral_size)
        p.pts(p1, s1)
        want_p.append(host(p1))
        want_s.append(host(s1))
    p2 = [empty(p.hybrid_size) for p in plans]
    s2 = [empty(p.spectral_size) for p in plans]
    for c in range(4):
        batch.set_schedule(stp=c, pts=c)
        batch.stp(s, p2)
        batch.pts(p2, s2)
        for i in range(len(plans)):
            assert np.array_equal(host(p2[i]), want_p[i]), (c, i)
            assert np.array_equal(host(s2[i]), want_s[i]), (c, i)
    times = batch.autotune(s, p2, s2, reps=2)
    assert set(times["stp"]) == {0, 2}
    assert set(times["pts"]) == ({0, 1, 2, 3} if precision == "fp64" else {0, 2})
    chosen = batch.schedule()
    assert chosen["stp"] == min(times["stp"], key=times["stp"].get)
    assert chosen["pts"] == min(times["pts"], key=times["pts"].get)
    batch.stp(s, p2)
    batch.pts(p2, s2)
    for i in range(len(plans)):
        assert np.array_equal(host(p2[i]), want_p[i]) and np.array_equal(host(s2[i]), want_s[i]), i
    from rayleigh_amd._lib import RayshtError
    with pytest.raises(RayshtError):
        batch.set_schedule(stp=4)


def test_batch_graph_replays(ra):
    """The batch round trip in a hipGraph: each replay, with eager work and
    synchronisations between replays, equals the eager grouped launches."""
    import torch
    plans, inputs = make(ra, "fp64")
    batch = ra.LegendreBatch(plans)
    s = [dev(x) for x in inputs]
    p2 = [empty(p.hybrid_size) for p in plans]
    eager = [empty(p.spectral_size) for p in plans]
    batch.stp(s, p2)
    batch.pts(p2, eager)
    want = [host(e) for e in eager]
    for i, (lmax, nt, kind, nr, nf) in enumerate(PROBLEMS):
        if nt == 2 * ((3 * (lmax + 1)) // 4) and kind == "all" and lmax > 0:
            # full m set, N_theta >= 3(l_max+1)/2: N_phi PtS(StP(s)) = s with N_phi = 2 N_theta
            assert rel_max(want[i] * (2 * nt), inputs[i]) < 1e-11, i
    out = [empty(p.spectral_size) for p in plans]
    g = batch.round_trip_graph(s, p2, out)
    for rep in range(3):
        for o in out:
            o.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        for i in range(len(plans)):
            assert np.array_equal(host(out[i]), want[i]), f"replay {rep} problem {i}"
        batch.stp(s, p2)
        torch.cuda.synchronize()


def test_batch_rejects_bad_input(ra):
    from rayleigh_amd._lib import RayshtError
    a = ra.LegendrePlan(48, 31, np.arange(32, dtype=np.int32), 2, 1)
    b = ra.LegendrePlan(48, 31, np.arange(32, dtype=np.int32), 2, 1, precision="bf16x6")
    with pytest.raises(RayshtError):
        ra.LegendreBatch([a, b])  # mixed precision
    with pytest.raises(RayshtError):
        ra.LegendreBatch([])
    many = [ra.LegendrePlan(4, 1, np.arange(2, dtype=np.int32), 1, 1) for _ in range(65)]
    with pytest.raises(RayshtError):
        ra.LegendreBatch(many)
    batch = ra.LegendreBatch([a])
    with pytest.raises(ValueError):
        batch.stp([dev(np.zeros(a.spectral_size))], [])
