"""Pins of the AMG oracle (oracle/amg.py) on the CPU.

Each vectorised oracle function is checked against a per-row brute-force
statement of its definition, against the mathematics it must satisfy
(P reproduces constants, Galerkin = dense R A P, FGMRES = direct solve), and
against the reference's own known-answer test (12x12 anchored Laplacian,
tests/test_amg_pressure_solver.py:89-106)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from matrices import apply_reference_fix, impes_like, laplacian_2d
from oracle import amg as oa

REF = oa.Hierarchy.REFERENCE


def _small_mats():
    A1 = apply_reference_fix(laplacian_2d(7), 0)
    A2, *_ = impes_like(6, 5, 3, seed=3)
    A3, *_ = impes_like(9, 7, 1, seed=4, sigma=1.5)
    return [A1, A2, A3]


def _row(A, i):
    s, e = A.indptr[i], A.indptr[i + 1]
    return A.indices[s:e], A.data[s:e]


def _brute_strength(A, theta):
    out = []
    for i in range(A.shape[0]):
        c, v = _row(A, i)
        d = abs(v[c == i][0]) if np.any(c == i) else 0.0
        neg = [(-x / (d + 1e-30)) if (j != i and x / (d + 1e-30) < 0) else 0.0 for j, x in zip(c, v)]
        m = max(neg) if neg else 0.0
        out += [bool(j != i and x / (d + 1e-30) < 0 and ng >= theta * m) for j, x, ng in zip(c, v, neg)]
    return np.array(out)


@pytest.mark.parametrize("theta", [0.25, 0.5])
def test_strength_matches_brute_force(theta):
    for A in _small_mats():
        assert np.array_equal(oa.strength(A, theta), _brute_strength(A, theta))


def test_cf_split_is_strict_local_max_plus_lift():
    for A in _small_mats():
        s = oa.strength(A, 0.25)
        cf, nc = oa.cf_split(A, s)
        n = A.shape[0]
        S = sp.csr_matrix((s.astype(float), A.indices, A.indptr), shape=A.shape)
        G = ((S + S.T) > 0).tolil()
        G.setdiag(0)
        G = sp.csr_matrix(G)
        G.eliminate_zeros()
        deg = np.diff(G.indptr)
        h = oa.hash32(np.arange(n)).astype(np.int64)
        key = [(int(deg[i]), int(h[i]), i) for i in range(n)]
        is_max = np.array([all(key[i] > key[j] for j in G.indices[G.indptr[i]:G.indptr[i + 1]]) for i in range(n)])
        has_c = np.array([any(is_max[j] for j in G.indices[G.indptr[i]:G.indptr[i + 1]]) for i in range(n)])
        want = np.where(is_max | ~has_c, oa.C_PT, oa.F_PT)
        assert np.array_equal(cf, want)
        assert nc == int(np.sum(want == oa.C_PT))
        # no two strong-graph neighbours are both local maxima
        r = np.repeat(np.arange(n), np.diff(G.indptr))
        assert not np.any(is_max[r] & is_max[G.indices])


def test_iterated_split_is_a_maximal_independent_set_built_round_by_round():
    for A in _small_mats() + [impes_like(14, 12, 5, seed=8, sigma=1.5)[0]]:
        s = oa.strength(A, 0.25)
        cf, nc = oa.cf_split_iterated(A, s)
        n = A.shape[0]
        S = sp.csr_matrix((s.astype(float), A.indices, A.indptr), shape=A.shape)
        G = ((S + S.T) > 0).tolil()
        G.setdiag(0)
        G = sp.csr_matrix(G)
        G.eliminate_zeros()
        nb = [G.indices[G.indptr[i]:G.indptr[i + 1]] for i in range(n)]
        deg = np.diff(G.indptr)
        h = oa.hash32(np.arange(n)).astype(np.int64)
        key = [(int(deg[i]), int(h[i]), i) for i in range(n)]
        # sequential restatement: one round at a time on python sets
        und, want = set(range(n)), {}
        while und:
            newc = {i for i in und if all(key[i] > key[j] for j in nb[i] if j in und)}
            assert newc  # progress in every round
            for i in newc:
                want[i] = oa.C_PT
            und -= newc
            cov = {i for i in und if any(j in newc for j in nb[i])}
            for i in cov:
                want[i] = oa.F_PT
            und -= cov
        assert np.array_equal(cf, np.array([want[i] for i in range(n)]))
        assert nc == int(np.sum(cf == oa.C_PT))
        isC = cf == oa.C_PT
        for i in range(n):
            if isC[i]:
                assert not np.any(isC[nb[i]])  # independent
            else:
                assert np.any(isC[nb[i]])      # maximal


def test_orphan_promotion_leaves_every_f_point_a_strong_c_neighbour():
    for A in _small_mats():
        s = oa.strength(A, 0.25)
        cf0, _ = oa.cf_split(A, s)
        cf = oa.promote_orphans(A, s, cf0)
        assert np.all(cf[cf0 == oa.C_PT] == oa.C_PT)
        for i in np.nonzero(cf == oa.F_PT)[0]:
            c, _ = _row(A, i)
            m = s[A.indptr[i]:A.indptr[i + 1]] & (c != i)
            assert np.any(cf[c[m]] == oa.C_PT)


def _brute_interp(A, cf, s):
    """Per-row statement of classical_amg.py:191-346 (RS direct + distance-two,
    weak-sum correction, clamp, F-row normalisation, zero-row fallback)."""
    n = A.shape[0]
    isC = cf == oa.C_PT
    f2c = np.cumsum(isC) - 1
    diag = np.array([(_row(A, i)[1][_row(A, i)[0] == i][0] if np.any(_row(A, i)[0] == i) else 0.0)
                     for i in range(n)])
    diag = np.where(np.abs(diag) > 1e-14, diag, 1.0)

    def strong_c(i):
        c, v = _row(A, i)
        st = s[A.indptr[i]:A.indptr[i + 1]]
        return [(j, x) for j, x, t in zip(c, v, st) if t and j != i and isC[j]]

    ssum = np.array([sum(-x for _, x in strong_c(i)) for i in range(n)])
    valid = np.array([len(strong_c(i)) > 0 and ssum[i] > 1e-14 for i in range(n)])
    P = np.zeros((n, max(int(isC.sum()), 1)))
    for i in range(n):
        if isC[i]:
            P[i, f2c[i]] = 1.0
            continue
        c, v = _row(A, i)
        st = s[A.indptr[i]:A.indptr[i + 1]]
        weak_c = sum(-x for j, x, t in zip(c, v, st) if j != i and not t and isC[j])
        f_weak = sum(-x for j, x, t in zip(c, v, st) if j != i and not isC[j] and (not t or not valid[j]))
        corr = (weak_c + f_weak) / ssum[i] if ssum[i] > 1e-14 else 0.0
        for j, x in strong_c(i):
            P[i, f2c[j]] += max((-x / diag[i]) * (1.0 + corr), 0.0)
        for j, x, t in zip(c, v, st):
            if t and j != i and not isC[j] and valid[j]:
                fac = -x / diag[i]
                for k, y in strong_c(j):
                    P[i, f2c[k]] += max(fac * (-y / ssum[j]), 0.0)
        rs = P[i].sum()
        if rs > 1e-14:
            P[i] /= rs
        else:
            cn = [j for j in c if isC[j]]
            if cn:
                P[i, f2c[cn[0]]] += 1.0
                P[i] /= P[i].sum()
    return np.maximum(P, 0.0)


def test_interpolation_matches_brute_force_and_reproduces_constants():
    for A in _small_mats():
        s = oa.strength(A, 0.25)
        cf0, _ = oa.cf_split(A, s)
        cf = oa.promote_orphans(A, s, cf0)
        P = oa.interpolation(A, cf, s)
        Pb = _brute_interp(A, cf, s)
        assert np.allclose(P.toarray(), Pb, rtol=1e-12, atol=1e-15)
        assert P.data.min() > 0.0  # explicit zeros dropped, weights clamped
        np.testing.assert_allclose(P @ np.ones(P.shape[1]), 1.0, rtol=1e-12)
        C = np.nonzero(cf == oa.C_PT)[0]
        assert np.array_equal(P[C].toarray(), np.eye(P.shape[1]))


def test_truncation_definition():
    rng = np.random.default_rng(0)
    n, nc = 40, 12
    dense = rng.random((n, nc)) * (rng.random((n, nc)) < 0.6)
    dense[3, :] = 0.25  # all ties
    dense[5, :] = 0.0
    dense[5, 7] = 1.0   # single weight: untouched
    cf = np.full(n, oa.F_PT, dtype=np.int8)
    cf[:2] = oa.C_PT
    P = sp.csr_matrix(dense)
    T = oa.truncate(P, cf, 0.2, 4).toarray()
    for i in range(n):
        row = dense[i]
        if cf[i] == oa.C_PT or np.count_nonzero(row) <= 1:
            assert np.array_equal(T[i], row)
            continue
        keep = (row > 0) & (row >= 0.2 * row.max())
        idx = np.nonzero(keep)[0]
        if idx.size > 4:  # four largest, ties to the lower column
            idx = np.sort(sorted(idx, key=lambda j: (-row[j], j))[:4])
        want = np.zeros(nc)
        want[idx] = row[idx] * row.sum() / row[idx].sum()
        np.testing.assert_allclose(T[i], want, rtol=1e-14, atol=0)
        assert np.isclose(T[i].sum(), row.sum(), rtol=1e-14)
    assert np.count_nonzero(T[3]) == 4 and np.array_equal(np.nonzero(T[3])[0], [0, 1, 2, 3])
    assert oa.truncate(P, cf, 0.0, 0) is P


@pytest.mark.parametrize("mode", ["reference", "default"])
def test_galerkin_operators_are_dense_rap(mode):
    A, *_ = impes_like(12, 10, 4, seed=9)
    kw = dict(REF) if mode == "reference" else {}
    H = oa.Hierarchy(A, coarsest_size=50, **kw)
    assert len(H.levels) > 2
    for l in range(len(H.levels) - 1):
        L, Lc = H.levels[l], H.levels[l + 1]
        Ad, Pd = L["A"].toarray(), L["P"].toarray()
        want = Pd.T @ Ad @ Pd
        assert np.allclose(Lc["A"].toarray(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        np.testing.assert_allclose(L["P"] @ np.ones(Pd.shape[1]), 1.0, rtol=1e-12)
        if mode == "default":
            cnt = np.diff(L["P"].indptr)
            assert cnt.max() <= 4


@pytest.mark.parametrize("mode", ["reference", "default"])
def test_reference_kat_anchored_laplacian(mode):
    """tests/test_amg_pressure_solver.py:89-106 of the reference: three cycles of
    apply() on the anchored 12x12 Laplacian reach ||Az - v|| / ||v|| < 5e-2 for a
    constant, a ramp and a random right-hand side."""
    A = apply_reference_fix(laplacian_2d(12), 0)
    kw = dict(REF) if mode == "reference" else {}
    H = oa.Hierarchy(A, theta=0.25, max_levels=10, coarsest_size=40, **kw)
    n = A.shape[0]
    vs = [np.ones(n), np.linspace(0, 1, n), np.random.default_rng(0).standard_normal(n)]
    for v in vs:
        v = v / np.linalg.norm(v)
        z = H.apply(v, cycles=3)
        assert np.linalg.norm(A @ z - v) / np.linalg.norm(v) < 5e-2


@pytest.mark.parametrize("mode", ["reference", "default"])
def test_fgmres_matches_direct_solve(mode):
    A, *_ = impes_like(20, 20, 5, seed=11)
    b = A @ (2e7 + 1e5 * np.random.default_rng(4).standard_normal(A.shape[0]))
    kw = dict(REF) if mode == "reference" else {}
    H = oa.Hierarchy(A, **kw)
    x, st = oa.fgmres(A, b, precond=H.precondition, rtol=1e-12)
    assert st["converged"]
    xs = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_scalar_scaling_is_invariant_to_matrix_scale():
    A, *_ = impes_like(14, 12, 3, seed=2)
    H1 = oa.Hierarchy(A)
    H2 = oa.Hierarchy(A * 1e-3)
    assert H1.equilibrated and H2.equilibrated
    d = np.abs(H1.levels[0]["A"].diagonal())
    assert np.isclose(np.sort(d)[(d.size - 1) // 2], 1.0, rtol=1e-14)  # lower median -> 1
    assert len(H1.levels) == len(H2.levels)
    for L1, L2 in zip(H1.levels[:-1], H2.levels[:-1]):
        assert np.array_equal(L1["cf"], L2["cf"])
        assert np.allclose(L1["P"].toarray(), L2["P"].toarray(), rtol=1e-10, atol=1e-14)
    v = np.random.default_rng(1).standard_normal(A.shape[0])
    # the preconditioner of D^-1 A is unchanged when A (and D) are scaled
    np.testing.assert_allclose(H2.precondition(v), H1.precondition(v), rtol=1e-9)
    # the scale is frozen at setup: an update with scaled values keeps it
    s0 = H1.scalar
    H1.update(A * 2.0)
    assert H1.scalar == s0


def test_coarse_pin_decision():
    # pure Neumann Laplacian: singular -> auto pins
    A = laplacian_2d(20)
    H = oa.Hierarchy(A, coarsest_size=40)
    assert len(H.levels) > 1 and H.levels[-1]["pin"]
    # nonsingular, nearly singular (small accumulation): auto does not pin, reference does
    B, *_ = impes_like(20, 20, 2, seed=1)
    assert not oa.Hierarchy(B).levels[-1]["pin"]
    assert oa.Hierarchy(B, coarse_pin=1).levels[-1]["pin"]
    assert not oa.Hierarchy(B, coarse_pin=0).levels[-1]["pin"]
    # single level: never pinned
    Hs = oa.Hierarchy(laplacian_2d(5), coarsest_size=200)
    assert len(Hs.levels) == 1 and not Hs.levels[0]["pin"]


def test_heterogeneous_nearly_singular_system_converges_in_few_iterations():
    """Lognormal (sigma 1.5) 7-point operator whose row sums are cut to 1% of the
    IMPES-like accumulation: the symmetric equilibration of the reference moves
    the near-nullspace to D^1/2 * 1, which the row-normalised interpolation does
    not reproduce; the scalar scaling keeps it at the constant vector."""
    A, *_ = impes_like(16, 16, 8, seed=3, wells=2, sigma=1.5)
    rs = np.asarray(A.sum(axis=1)).ravel()
    A = A.tolil()
    A.setdiag(A.diagonal() - 0.99 * rs)
    A = A.tocsr()
    b = A @ np.random.default_rng(0).standard_normal(A.shape[0])
    H = oa.Hierarchy(A, max_levels=30)
    x, st = oa.fgmres(A, b, precond=H.precondition, rtol=1e-10)
    assert st["converged"] and st["iterations"] <= 20, st
    Hr = oa.Hierarchy(A, max_levels=30, **REF)
    xr, sr = oa.fgmres(A, b, precond=Hr.precondition, rtol=1e-10, maxit=300)
    assert sr["iterations"] > 3 * st["iterations"], (st, sr)
    xs = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_dominant_rows_become_coarse_points():
    """On an assembled IMPES pressure matrix (the 200k-cell case shrunk to
    16x16x4, with its BHP producer), the BHP well row (|a_ii| ~ 1e16 x its
    couplings) is a C point and the coarsest operator stays well conditioned
    after diagonal scaling. Without the rule the well row is an F point whose
    row-normalised interpolation spreads its diagonal over several coarse
    unknowns: the scaled coarsest operator is then singular to ~1e-13."""
    from oracle import impes as oi
    from simulator.cases import baseline_case

    cfg, f = baseline_case(1, dims=(16, 16, 4))
    case = oi.Case(cfg, f, pvt=oi.PVT(cfg["fluid"]["pvt_table"]))
    ps = oi.pressure_system(case, case.initial_state(), cfg["simulation"]["time_step_days"] * 86400)
    A = oi.system_csr(case, ps)
    d = A.diagonal()
    wells = np.nonzero(d > 1e6 * np.median(d))[0]
    assert wells.size == 1

    def coarsest_min_eig(H):
        C = H.levels[-1]["A"].toarray()
        dg = np.abs(np.diag(C))
        Cs = C / np.sqrt(np.outer(dg, dg))
        return np.linalg.eigvalsh((Cs + Cs.T) / 2).min()

    H = oa.Hierarchy(A, coarsest_size=40)
    assert np.all(H.levels[0]["cf"][wells] == oa.C_PT)
    assert coarsest_min_eig(H) > 1e-3
    Hr = oa.Hierarchy(A, coarsest_size=40, dominant=0.0)
    assert coarsest_min_eig(Hr) < 1e-10
    b = ps["b"]
    _, st = oa.fgmres(A, b, precond=H.precondition, rtol=1e-10)
    _, sr = oa.fgmres(A, b, precond=Hr.precondition, rtol=1e-10)
    assert st["iterations"] <= sr["iterations"]
