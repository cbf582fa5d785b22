"""GPU parity of the solver kernels against the CPU oracle (oracle/amg.py) and scipy.

Bars: integer data (C/F markers, sparsity patterns) bit-exact; FP64 values to
1e-12 relative (summation order differs); FP32 V-cycle to 1e-5 relative;
FGMRES solutions to 1e-8 relative against scipy's direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

from matrices import apply_reference_fix, impes_like, laplacian_2d
from oracle import amg as oa

pytestmark = pytest.mark.gpu


def to_dev(A: sp.csr_matrix, dev, dtype=torch.float64):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return (torch.tensor(A.indptr, dtype=torch.int32, device=dev),
            torch.tensor(A.indices, dtype=torch.int32, device=dev),
            torch.tensor(A.data, dtype=dtype, device=dev))


def t(x, dev, dtype=torch.float64):
    return torch.tensor(np.asarray(x), dtype=dtype, device=dev)


def csr_from_export(e, ncols=None):
    n = e["n"]
    A = sp.csr_matrix((e["val"].cpu().numpy(), e["col"].cpu().numpy(), e["rowptr"].cpu().numpy()), shape=(n, n))
    return A


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dims", [(6, 5, 4), (20, 20, 10), (33, 17, 1)])
def test_spmv_csr_and_stencil(dev, dims):
    import impes_hip as H

    A, d, ox, oy, oz = impes_like(*dims, seed=3)
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    ref = A @ x
    rp, ci, v = to_dev(A, dev)
    y = H.spmv_csr(rp, ci, v, t(x, dev)).cpu().numpy()
    assert rel(y, ref) < 1e-14
    y32 = H.spmv_csr(rp, ci, v.float(), t(x, dev, torch.float32)).cpu().numpy()
    assert rel(y32, ref) < 1e-5
    ys = H.spmv_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev), t(x, dev)).cpu().numpy()
    assert rel(ys, ref) < 1e-14


@pytest.mark.parametrize("dims", [(6, 5, 4), (7, 3, 8), (4, 4, 2), (5, 6, 10)])
def test_spmv_st7_wide_equals_scalar(dev, dims):
    """Stencil products over whole 16-byte vectors along z (even nz in FP64)
    against the one-cell-per-thread form, selected here by an x one element off
    16-byte alignment: bit-equal, and equal to the CSR product to 1e-14."""
    import impes_hip as H

    A, d, ox, oy, oz = impes_like(*dims, seed=11)
    x = np.random.default_rng(6).standard_normal(A.shape[0])
    xa = t(x, dev)
    buf = torch.empty(A.shape[0] + 1, dtype=torch.float64, device=dev)
    xu = buf[1:]
    xu.copy_(xa)
    assert xa.data_ptr() % 16 == 0 and xu.data_ptr() % 16 != 0
    args = (dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    yw = H.spmv_st7(*args, xa)
    ys = H.spmv_st7(*args, xu)
    assert torch.equal(yw, ys)
    assert rel(yw.cpu().numpy(), A @ x) < 1e-14


def test_l1jacobi_and_dot(dev):
    import impes_hip as H

    A, *_ = impes_like(10, 9, 8, seed=4)
    rng = np.random.default_rng(1)
    x, b = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    ir = oa.inv_relax(A)
    ref = x + 0.7 * ir * (b - A @ x)
    rp, ci, v = to_dev(A, dev)
    out = H.l1jacobi_csr(rp, ci, v, t(ir, dev), 0.7, t(x, dev), t(b, dev)).cpu().numpy()
    assert rel(out, ref) < 1e-13
    dd = H.dot(t(x, dev), t(b, dev)).item()
    assert abs(dd - x @ b) <= 1e-12 * np.abs(x * b).sum()


def _compare_hierarchy(hg, ho, tol=1e-11):
    assert hg.num_levels == len(ho.levels)
    assert hg.equilibrated == ho.equilibrated
    for l, L in enumerate(ho.levels):
        e = hg.export_level(l)
        Ag = csr_from_export(e)
        Ao = L["A"]
        assert e["n"] == Ao.shape[0]
        assert bool(e["is_coarsest"]) == L["coarsest"]
        diff = Ag - Ao
        scale = abs(Ao).max()
        assert (abs(diff).max() if diff.nnz else 0.0) <= tol * scale, f"level {l} operator"
        assert rel(e["inv_relax"].cpu().numpy(), L["ir"]) < tol
        if not L["coarsest"]:
            cf = e["cf"].cpu().numpy()
            assert np.array_equal(cf, L["cf"]), f"level {l} C/F split"
            P = sp.csr_matrix((e["p_val"].cpu().numpy(), e["p_col"].cpu().numpy(), e["p_rowptr"].cpu().numpy()),
                              shape=L["P"].shape)
            assert np.array_equal(P.indptr, L["P"].indptr), f"level {l} P pattern"
            assert np.array_equal(P.indices, L["P"].indices), f"level {l} P pattern"
            assert np.max(np.abs(P.data - L["P"].data)) < 1e-13, f"level {l} P values"


@pytest.mark.parametrize("mode", ["reference", "default"])
@pytest.mark.parametrize("case", ["laplacian", "neumann", "impes_2d", "impes_3d"])
def test_hierarchy_matches_oracle(dev, case, mode):
    import impes_hip as H

    if case == "laplacian":
        A = apply_reference_fix(laplacian_2d(12), 0)
        kw = dict(theta=0.25, max_levels=10, coarsest_size=40)
    elif case == "neumann":  # singular: the coarsest level is pinned in both modes
        A = laplacian_2d(24)
        kw = dict(theta=0.25, max_levels=10, coarsest_size=60)
    elif case == "impes_2d":
        A, *_ = impes_like(60, 60, 1, seed=1)
        kw = dict(theta=0.25, max_levels=10, coarsest_size=200)
    else:
        A, *_ = impes_like(20, 20, 10, seed=1)
        kw = dict(theta=0.25, max_levels=10, coarsest_size=200)
    if mode == "reference":
        kw.update(oa.Hierarchy.REFERENCE)
    ho = oa.Hierarchy(A, **kw)
    hg = H.AmgHierarchy(**kw)
    hg.setup_csr(*to_dev(A, dev))
    _compare_hierarchy(hg, ho)
    # V-cycle parity (equilibrated space), FP64 and FP32
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    xo = ho.vcycle(b)
    xg = hg.vcycle(t(b, dev), precision=64).cpu().numpy()
    assert rel(xg, xo) < 1e-11
    xg32 = hg.vcycle(t(b, dev), precision=32).cpu().numpy()
    assert rel(xg32, xo) < 1e-4


def test_padded_sweep_rows_longer_than_a_wave_tile(dev):
    """A level with a row of more than 64 four-entry chunks (a hub cell coupled
    to 400 others) cannot be cut into wave tiles of whole rows: that level's
    FP32 sweeps take the lanes-per-row kernel, the others the tiled one. The
    hierarchy and both V-cycles still match the oracle."""
    import impes_hip as H

    A, *_ = impes_like(16, 16, 6, seed=3)
    A = sp.lil_matrix(A)
    hub = np.random.default_rng(4).choice(np.arange(1, A.shape[0]), 400, replace=False)
    c = 0.05 * abs(A[0, 0])
    for j in hub:
        A[0, j] = A[j, 0] = -c
        A[j, j] = A[j, j] + c
    A[0, 0] = A[0, 0] + c * len(hub)
    A = sp.csr_matrix(A)
    assert np.diff(A.indptr).max() > 256
    kw = dict(theta=0.25, max_levels=10, coarsest_size=100)
    ho = oa.Hierarchy(A, **kw)
    hg = H.AmgHierarchy(**kw)
    hg.setup_csr(*to_dev(A, dev))
    _compare_hierarchy(hg, ho)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    xo = ho.vcycle(b)
    assert rel(hg.vcycle(t(b, dev), precision=64).cpu().numpy(), xo) < 1e-11
    assert rel(hg.vcycle(t(b, dev), precision=32).cpu().numpy(), xo) < 1e-4


@pytest.mark.parametrize("case", ["max_levels_1", "max_levels_2", "single_level_swept", "single_level_dense",
                                  "single_level_dense_negdef", "mfma_2342", "mfma_695", "mfma_1346_pinned",
                                  "mfma_single_4000"])
def test_coarsest_level_edge_cases(dev, case):
    """Coarsest levels above the dense limit (500 rows: solved by 50 L1-Jacobi
    sweeps, classical_amg.py:1045-1069) after one or two coarsenings, and
    one-level hierarchies (no pin) solved by sweeps or densely (400 rows: the
    matrix-core inverse, and for the negated, negative definite operator its
    fallback to pivoted Gauss-Jordan); dense coarsest levels of 695 to 4000 rows
    (dense_max 4096), whose inverse is formed on the matrix cores (blocked
    Gauss-Jordan, MFMA FP64), with and without the DOF-0 pin (the reference's
    hierarchy pins unconditionally): hierarchy and V-cycle against the oracle
    (numpy's dense inverse)."""
    import impes_hip as H

    mfma = {"mfma_2342": ((40, 40, 10), dict(coarsest_size=3000, dense_max=4096)),
            "mfma_695": ((40, 40, 10), dict(coarsest_size=1500, dense_max=4096)),
            "mfma_1346_pinned": ((30, 30, 12), dict(coarsest_size=1500, dense_max=4096, **oa.Hierarchy.REFERENCE)),
            "mfma_single_4000": ((20, 20, 10), dict(coarsest_size=5000, dense_max=4096))}
    if case in ("single_level_dense", "single_level_dense_negdef"):
        A, *_ = impes_like(10, 8, 5, seed=4)
        if case == "single_level_dense_negdef":  # not SPD: the matrix-core inverse falls back to pivoting
            A = -A
        kw = dict(coarsest_size=1000)
    elif case in mfma:
        dims, kw = mfma[case]
        A, *_ = impes_like(*dims, seed=4)
    else:
        A, *_ = impes_like(20, 20, 10, seed=4)
        kw = {"max_levels_1": dict(max_levels=1), "max_levels_2": dict(max_levels=2),
              "single_level_swept": dict(coarsest_size=5000)}[case]
    ho = oa.Hierarchy(A, **kw)
    hg = H.AmgHierarchy(**kw)
    hg.setup_csr(*to_dev(A, dev))
    _compare_hierarchy(hg, ho)
    info = hg.level_info(hg.num_levels - 1)
    assert bool(info["dense"]) == (case.startswith("single_level_dense") or case in mfma)
    if case in mfma:
        assert info["n"] > 512 and ho.levels[-1]["pin"] == (case == "mfma_1346_pinned")
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    xo = ho.vcycle(b)
    assert rel(hg.vcycle(t(b, dev), precision=64).cpu().numpy(), xo) < 1e-11
    assert rel(hg.vcycle(t(b, dev), precision=32).cpu().numpy(), xo) < 1e-4


@pytest.mark.parametrize("setup", ["csr", "stencil"])
def test_level_times_cover_every_level(dev, setup):
    """impes_amg_level_times: a positive, finite time and byte count for every
    level in both precisions, bytes at least the level's four operator passes
    (V(2,2) with the first sweep fused into the restriction above it), a positive
    graph-replayed cycle time, and the preconditioner unchanged by the timing
    cycles."""
    import impes_hip as H

    dims = (20, 20, 10)
    A, d, ox, oy, oz = impes_like(*dims, seed=1)
    hg = H.AmgHierarchy()
    if setup == "csr":
        hg.setup_csr(*to_dev(A, dev))
    else:
        hg.setup_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    v = t(np.random.default_rng(3).standard_normal(A.shape[0]), dev)
    z0 = hg.precondition(v, precision=32).clone()
    for prec in (32, 64):
        ms, by, cyc = hg.level_times(precision=prec, cycles=3)
        assert len(ms) == len(by) == hg.num_levels
        assert all(np.isfinite(m) and m > 0 for m in ms) and cyc > 0
        vb = 4 if prec == 32 else 8
        for l in range(1, hg.num_levels - 1):
            assert by[l] >= 4 * hg.level_info(l)["nnz"] * (4 + vb), l
        assert all(b > 0 for b in by)
    assert torch.equal(hg.precondition(v, precision=32), z0)


def test_hierarchy_matches_oracle_131k_rows(dev):
    """131k rows: coarse levels with 40-70 entries per row (the setup's
    wavefront-per-row kernels) and product lists of up to ~4200 entries per row
    (every LDS bin of the segmented sort)."""
    import impes_hip as H

    A, *_ = impes_like(64, 64, 32, seed=3)
    ho = oa.Hierarchy(A)
    hg = H.AmgHierarchy()
    hg.setup_csr(*to_dev(A, dev))
    _compare_hierarchy(hg, ho)
    b = np.random.default_rng(10).standard_normal(A.shape[0])
    assert rel(hg.vcycle(t(b, dev), precision=64).cpu().numpy(), ho.vcycle(b)) < 1e-11


def test_galerkin_identity_at_1m_rows(dev):
    """At 1M rows (product lists beyond 8192 entries per row: the global-scratch
    sort), checked without the oracle: every exported level has strictly
    increasing columns per row, and A_{l+1} = P^T A_l P to rounding, recomputed
    with scipy from the exported A_l and P."""
    import impes_hip as H

    A, *_ = impes_like(100, 100, 100, seed=3)
    hg = H.AmgHierarchy()
    hg.setup_csr(*to_dev(A, dev))
    assert hg.num_levels >= 6
    prev = None
    for l in range(hg.num_levels):
        e = hg.export_level(l)
        Al = csr_from_export(e)
        d = np.diff(Al.indices)
        assert np.all((d > 0) | (np.diff(np.repeat(np.arange(Al.shape[0]), np.diff(Al.indptr))) != 0)), l
        if prev is not None:
            Ap, P = prev
            G = (P.T.tocsr() @ (Ap @ P)).tocsr()
            diff = G - Al
            assert (abs(diff).max() if diff.nnz else 0.0) <= 1e-11 * abs(Al).max(), f"level {l} Galerkin"
        if not e["is_coarsest"]:
            P = sp.csr_matrix((e["p_val"].cpu().numpy(), e["p_col"].cpu().numpy(), e["p_rowptr"].cpu().numpy()),
                              shape=(e["n"], hg.level_info(l + 1)["n"]))
            dp = np.diff(P.indices)
            assert np.all((dp > 0) | (np.diff(np.repeat(np.arange(P.shape[0]), np.diff(P.indptr))) != 0)), l
            prev = (Al, P)


def test_hierarchy_update_values(dev):
    import impes_hip as H

    # same pattern and well cells (same seed), new transmissibilities: a simulator's
    # next step (its wells do not move between steps)
    A1, *_ = impes_like(20, 20, 10, seed=5)
    A2, *_ = impes_like(20, 20, 10, seed=5, sigma=1.3)
    ho = oa.Hierarchy(A1)
    hg = H.AmgHierarchy()
    hg.setup_csr(*to_dev(A1, dev))
    ho.update(A2)
    hg.update_csr(to_dev(A2, dev)[2])
    _compare_hierarchy(hg, ho)


def test_repeated_setup_on_one_handle(dev):
    """A handle set up again (new pattern) equals a fresh handle, and rebuilds do
    not accumulate device memory: each build's blocks come from the handle's
    setup pool, and the previous build's blocks are freed when the next build ends."""
    import impes_hip as H

    mats = [impes_like(20, 20, 10, seed=s)[0] for s in (6, 7)] + [impes_like(24, 16, 8, seed=8)[0]]
    b = {A.shape[0]: np.random.default_rng(9).standard_normal(A.shape[0]) for A in mats}
    hg = H.AmgHierarchy()
    free = []
    for rep in range(3):
        for A in mats:
            hg.setup_csr(*to_dev(A, dev))
            fresh = H.AmgHierarchy()
            fresh.setup_csr(*to_dev(A, dev))
            x = hg.vcycle(t(b[A.shape[0]], dev), precision=64)
            y = fresh.vcycle(t(b[A.shape[0]], dev), precision=64)
            assert torch.equal(x, y)
            del fresh
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
    # the third round holds no more than the second (32 MiB slab granularity)
    assert free[2] >= free[1] - (64 << 20)


@pytest.mark.parametrize("dims", [(16, 12, 6), (12, 10, 8)])
def test_stencil_hierarchy_equals_csr_hierarchy(dev, dims):
    """nz = 6 runs the FP32 level-0 sweeps one cell per thread, nz = 8 four cells
    per thread with 16-byte loads (FP64: two cells per thread in both)."""
    import impes_hip as H

    A, d, ox, oy, oz = impes_like(*dims, seed=7)
    ho = oa.Hierarchy(A)
    hg = H.AmgHierarchy()
    hg.setup_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    _compare_hierarchy(hg, ho)
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    xo = ho.vcycle(b)
    assert rel(hg.vcycle(t(b, dev)).cpu().numpy(), xo) < 1e-11
    assert rel(hg.vcycle(t(b, dev), precision=32).cpu().numpy(), xo) < 1e-4


@pytest.mark.parametrize("mode", ["reference", "default"])
@pytest.mark.parametrize("coarsest", [120, 200])
def test_stencil_update_equals_csr_update(dev, mode, coarsest):
    """A numeric update from a new stencil (one fused pass over level 0) against
    the oracle's update and, exactly, against the CSR hierarchy's update of the
    same matrix: every level operator and every relaxation weight bit-equal.
    The coarsest sizes (53/116/156/160 rows) run each register tile of the
    coarsest inverse; scalar and diagonal equilibration both scale."""
    import impes_hip as H

    dims = (16, 12, 6)
    A1, d1, ox1, oy1, oz1 = impes_like(*dims, seed=9)
    A2, d2, ox2, oy2, oz2 = impes_like(*dims, seed=9, sigma=1.3)  # same well cells
    kw = dict(coarsest_size=coarsest)
    if mode == "reference":
        kw.update(oa.Hierarchy.REFERENCE)
    ho = oa.Hierarchy(A1, **kw)
    ho.update(A2)
    hs = H.AmgHierarchy(**kw)
    hs.setup_st7(dims, t(d1, dev), t(ox1, dev), t(oy1, dev), t(oz1, dev))
    hs.update_st7(t(d2, dev), t(ox2, dev), t(oy2, dev), t(oz2, dev))
    hc = H.AmgHierarchy(**kw)
    hc.setup_csr(*to_dev(A1, dev))
    hc.update_csr(to_dev(A2, dev)[2])
    assert ho.equilibrated
    _compare_hierarchy(hs, ho)
    for l in range(hs.num_levels):
        es, ec = hs.export_level(l), hc.export_level(l)
        diff = csr_from_export(es) - csr_from_export(ec)
        assert diff.nnz == 0 or abs(diff).max() == 0.0, f"level {l} operator"
        assert torch.equal(es["inv_relax"], ec["inv_relax"]), f"level {l} relaxation weights"
    b = np.random.default_rng(4).standard_normal(A1.shape[0])
    assert rel(hs.vcycle(t(b, dev), precision=64).cpu().numpy(), ho.vcycle(b)) < 1e-11
    v = t(b, dev)
    zs, zc = hs.precondition(v, precision=64), hc.precondition(v, precision=64)
    assert rel(zs.cpu().numpy(), zc.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("mode", ["reference", "default"])
def test_transposed_vcycle_is_the_adjoint(dev, mode):
    """<x, B y> = <B^T x, y> for the zero-start cycle B and its transposed entry
    point (smoothing counts swapped), on a symmetric IMPES operator. The
    reference's V(3,2) is not self-adjoint (the plain cycle fails the identity);
    the default V(2,2) is, and its transposed cycle is the cycle itself."""
    import impes_hip as H

    A, *_ = impes_like(16, 12, 6, seed=3)
    assert abs(A - A.T).max() == 0.0
    kw = dict(oa.Hierarchy.REFERENCE) if mode == "reference" else {}
    hg = H.AmgHierarchy(**kw)
    hg.setup_csr(*to_dev(A, dev))
    rng = np.random.default_rng(12)
    x, y = t(rng.standard_normal(A.shape[0]), dev), t(rng.standard_normal(A.shape[0]), dev)
    By, Bx = hg.vcycle(y), hg.vcycle(x)
    Btx = hg.vcycle(x, transposed=True)
    lhs, rhs = float(x @ By), float(Btx @ y)
    scale = float(x.norm() * By.norm())
    assert abs(lhs - rhs) <= 1e-12 * scale
    if mode == "reference":
        assert abs(lhs - float(Bx @ y)) > 1e-6 * scale
    else:
        assert torch.equal(Btx, Bx)
    # the FP32 cycle satisfies the identity to its own rounding
    By32, Btx32 = hg.vcycle(y, precision=32), hg.vcycle(x, precision=32, transposed=True)
    assert abs(float(x @ By32) - float(Btx32 @ y)) <= 1e-5 * scale


@pytest.mark.parametrize("precision", [64, 32])
def test_graph_replay_matches_eager_vcycle(dev, precision):
    """The captured V-cycle graph against eager launches of the same kernels:
    several replays with eager work, host synchronisation and a numeric update of
    the hierarchy between them; every replay must be bit-identical."""
    import impes_hip as H

    dims = (24, 20, 6)
    A1, d1, ox1, oy1, oz1 = impes_like(*dims, seed=21)
    A2, d2, ox2, oy2, oz2 = impes_like(*dims, seed=21, sigma=1.3)  # same well cells
    hg = H.AmgHierarchy()
    hg.setup_st7(dims, t(d1, dev), t(ox1, dev), t(oy1, dev), t(oz1, dev))
    rng = np.random.default_rng(5)
    for rep in range(6):
        if rep == 3:
            hg.update_st7(t(d2, dev), t(ox2, dev), t(oy2, dev), t(oz2, dev))
        v = t(rng.standard_normal(A1.shape[0]), dev)
        hg.set_graphs(True)
        zg = hg.precondition(v, precision=precision).clone()
        torch.cuda.synchronize()
        _ = (v * 2.0).sum().item()  # unrelated eager work in between
        hg.set_graphs(False)
        ze = hg.precondition(v, precision=precision)
        torch.cuda.synchronize()
        assert torch.equal(zg, ze), rep
    # and the eager V-cycle matches the oracle's preconditioner on the updated matrix
    ho = oa.Hierarchy(A1)
    ho.update(A2)
    vv = rng.standard_normal(A1.shape[0])
    hg.set_graphs(True)
    z = hg.precondition(t(vv, dev), precision=precision).cpu().numpy()
    assert rel(z, ho.precondition(vv)) < (1e-11 if precision == 64 else 1e-4)


@pytest.mark.parametrize("restart", [30, 6])
def test_deferred_convergence_tests_change_nothing(dev, restart, monkeypatch):
    """FGMRES enqueues a step before testing the previous one while the estimates
    predict no stop. Whatever the deferral (predictive by default, none with
    IMPES_FGMRES_LOOKAHEAD=0, every test with =1), a speculative step after the
    stop is ignored: the same bits and iteration counts, also across restarts
    (restart 6)."""
    import impes_hip as H

    dims = (24, 20, 8)
    A, d, ox, oy, oz = impes_like(*dims, seed=12)
    b = A @ (2e7 + 1e5 * np.random.default_rng(6).standard_normal(A.shape[0]))
    hg = H.AmgHierarchy()
    hg.setup_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    out = {}
    for mode in (None, "0", "1"):
        if mode is None:
            monkeypatch.delenv("IMPES_FGMRES_LOOKAHEAD", raising=False)
        else:
            monkeypatch.setenv("IMPES_FGMRES_LOOKAHEAD", mode)
        solver = H.FgmresSolver()  # reads the switch when created
        x = torch.zeros(A.shape[0], dtype=torch.float64, device=dev)
        st = solver.solve_st7(hg, dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev), t(1.0 / d, dev), t(b, dev),
                              x, precision=32, rtol=1e-10, restart=restart, maxit=300)
        assert st["converged"], (mode, st)
        out[mode] = (x.cpu().numpy(), st["iterations"], st["restarts"])
    for mode in ("0", "1"):
        assert np.array_equal(out[mode][0], out[None][0]), mode
        assert out[mode][1:] == out[None][1:], (mode, out[mode][1:], out[None][1:])


@pytest.mark.parametrize("restart,precision", [(30, 32), (6, 32), (30, 64), (32, 32)])
def test_lagged_orthogonalisation_matches_cgs2(dev, restart, precision, monkeypatch):
    """The default Arnoldi step (reorthogonalisation lagged by one step: two
    passes over the basis) against CGS2's three passes (IMPES_FGMRES_CGS2=1):
    the same basis up to rounding, so the same iteration and restart counts and
    solutions within the solve tolerance of each other and of scipy's direct
    solve; also at the largest restart of the lagged form (32) and across
    restarts (6)."""
    import impes_hip as H

    dims = (24, 20, 8)
    A, d, ox, oy, oz = impes_like(*dims, seed=12)
    xt = 2e7 + 1e5 * np.random.default_rng(6).standard_normal(A.shape[0])
    b = A @ xt
    hg = H.AmgHierarchy()
    hg.setup_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("IMPES_FGMRES_CGS2", mode)
        solver = H.FgmresSolver()  # reads the switch when created
        x = torch.zeros(A.shape[0], dtype=torch.float64, device=dev)
        st = solver.solve_st7(hg, dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev), t(1.0 / d, dev), t(b, dev),
                              x, precision=precision, rtol=1e-10, restart=restart, maxit=300)
        assert st["converged"], (mode, st)
        out[mode] = (x.cpu().numpy(), st["iterations"], st["restarts"])
    assert out["0"][1:] == out["1"][1:], (out["0"][1:], out["1"][1:])
    xd = spl.spsolve(sp.csc_matrix(A), b)
    for mode in ("0", "1"):
        assert rel(out[mode][0], xd) < 1e-8, mode
    assert rel(out["0"][0], out["1"][0]) < 1e-8


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("dims", [(60, 60, 1), (20, 20, 10)])
def test_fgmres_matches_direct_solve(dev, precision, dims):
    import impes_hip as H

    A, d, ox, oy, oz = impes_like(*dims, seed=11)
    rng = np.random.default_rng(4)
    b = (A @ (2e7 + 1e5 * rng.standard_normal(A.shape[0])))
    xs = spl.spsolve(A.tocsc(), b)
    hg = H.AmgHierarchy()
    hg.setup_st7(dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev))
    solver = H.FgmresSolver()
    x = torch.zeros(A.shape[0], dtype=torch.float64, device=dev)
    dinv = t(1.0 / d, dev)
    st = solver.solve_st7(hg, dims, t(d, dev), t(ox, dev), t(oy, dev), t(oz, dev), dinv, t(b, dev), x,
                          precision=precision, rtol=1e-12, restart=30, maxit=300)
    assert st["converged"], st
    assert rel(x.cpu().numpy(), xs) < 1e-8, st
    # same system through the CSR entry point
    x2 = torch.zeros_like(x)
    rp, ci, v = to_dev(A, dev)
    st2 = solver.solve_csr(hg, rp, ci, v, dinv, t(b, dev), x2, precision=precision, rtol=1e-12)
    assert st2["converged"]
    assert rel(x2.cpu().numpy(), xs) < 1e-8
    # oracle FGMRES with the oracle hierarchy takes a similar number of iterations
    if precision == 64:
        ho = oa.Hierarchy(A)
        _, sto = oa.fgmres(A, b, precond=ho.precondition, rtol=1e-12)
        assert abs(sto["iterations"] - st["iterations"]) <= 2, (sto, st)
