"""Known-answer tests of the oracle's CUSP-semantics building blocks.

The only known-answer vectors the reference holds for this path are the vendored CUSP unit
tests (include/cusplibrary-0.5.1/testing/): they pin the DIA layout and the Krylov/monitor
stopping semantics the reference solver relies on.  Each test below transcribes one of them.
"""
import numpy as np
import pytest

from oracle import oracle


def test_poisson5pt_2x3_exact_matrix():
    # testing/poisson.cu:7-60 -- the 6x6 matrix of poisson5pt(2, 3)
    E = np.array([
        [4, -1, -1, 0, 0, 0],
        [-1, 4, 0, -1, 0, 0],
        [-1, 0, 4, -1, -1, 0],
        [0, -1, -1, 4, 0, -1],
        [0, 0, -1, 0, 4, -1],
        [0, 0, 0, -1, -1, 4]], dtype=float)
    off, val = oracle.poisson5pt(2, 3)
    assert list(off) == [-2, -1, 0, 1, 2]
    assert np.array_equal(oracle.dia_to_dense(off, val), E)


def test_poisson7pt_offsets_and_rowsums():
    # cusp/gallery/detail/poisson.inl:66-86 stencil order (0,0,-1),(0,-1,0),(-1,0,0),0,...
    off, val = oracle.poisson7pt(3, 4, 5)
    assert list(off) == [-12, -3, -1, 0, 1, 3, 12]
    A = oracle.dia_to_dense(off, val)
    assert np.array_equal(A, A.T)
    interior = (1 * 3 * 4) + 1 * 3 + 1        # cell (1,1,1)
    assert A[interior].sum() == 0.0


def test_dia_spmv_matches_dense():
    rng = np.random.default_rng(0)
    off, val = oracle.poisson7pt(4, 3, 5)
    val = val * rng.uniform(0.5, 2.0, val.shape)
    x = rng.standard_normal(val.shape[1])
    y = oracle.dia_spmv(off, val, x)
    assert np.allclose(y, oracle.dia_to_dense(off, val) @ x, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("solver", [oracle.dia_cg, oracle.dia_bicgstab, oracle.dia_pcg_jacobi])
def test_poisson5pt_10x10_converges(solver):
    # testing/cg.cu:9-29, testing/bicgstab.cu:9-29: b = 1, 20 iterations, rtol 1e-4
    off, val = oracle.poisson5pt(10, 10)
    b = np.ones(100)
    x, it, rn, conv = solver(off, val, b, np.zeros(100), 20, 1e-4)
    A = oracle.dia_to_dense(off, val)
    assert np.linalg.norm(b - A @ x) < 1e-4 * np.linalg.norm(b)
    assert conv and it <= 20


@pytest.mark.parametrize("solver", [oracle.dia_cg, oracle.dia_bicgstab, oracle.dia_pcg_jacobi])
def test_zero_residual_start_does_no_iteration(solver):
    # testing/cg.cu:33-61: M = diag(8, 4), x = 1, b = M x, monitor(b, 20, 0)
    off = np.array([0], np.int32)
    val = np.array([[8.0, 4.0]])
    x0 = np.ones(2)
    b = np.array([8.0, 4.0])
    x, it, rn, conv = solver(off, val, b, x0, 20, 0.0)
    assert conv and it == 0 and rn == 0.0
    assert np.array_equal(x, x0)


def test_monitor_tolerance_is_atol_plus_rtol_bnorm():
    # testing/monitor.cu:6-30: tolerance = atol + rtol*|b|; finished() is converged-or-limit
    off, val = oracle.poisson5pt(4, 4)
    b = np.ones(16)
    x0 = np.zeros(16)
    r0 = np.linalg.norm(b)                  # |r0| = |b| = 4
    # atol + rtol*|b| == |r0| exactly -> converged before the first iteration (<=)
    _, it, rn, conv = oracle.dia_cg(off, val, b, x0, 5, 0.5, r0 - 0.5 * r0)
    assert it == 0 and conv and rn == r0
    # just below -> at least one iteration
    _, it, _, _ = oracle.dia_cg(off, val, b, x0, 5, 0.5, 0.5 * r0 * (1 - 1e-12))
    assert it >= 1


def test_iteration_limit_is_exact():
    off, val = oracle.poisson5pt(10, 10)
    b = np.ones(100)
    for solver in (oracle.dia_cg, oracle.dia_pcg_jacobi, oracle.dia_bicgstab):
        _, it, _, conv = solver(off, val, b, np.zeros(100), 3, 1e-30)
        assert it == 3 and not conv


def test_pcg_jacobi_matches_dense_lu():
    rng = np.random.default_rng(1)
    off, val = oracle.poisson7pt(5, 4, 6)
    n = val.shape[1]
    # SPD: symmetric positive off-diagonal scaling + diagonal shift
    A = oracle.dia_to_dense(off, val)
    w = rng.uniform(0.2, 3.0, n)
    A = A * np.sqrt(np.outer(w, w))
    A += np.diag(rng.uniform(0.01, 1.0, n))
    dense_val = np.zeros_like(val)
    for d, o in enumerate(off):
        for r in range(n):
            c = r + o
            if 0 <= c < n:
                dense_val[d, r] = A[r, c]
    b = rng.standard_normal(n)
    x_lu = oracle.dense_solve(A, b)
    assert np.allclose(A @ x_lu, b, atol=1e-12)
    x, it, rn, conv = oracle.dia_pcg_jacobi(off, dense_val, b, np.zeros(n), 1000, 1e-13)
    assert conv
    assert np.max(np.abs(x - x_lu)) < 1e-10 * np.max(np.abs(x_lu))


def test_thomas_matches_dense():
    rng = np.random.default_rng(2)
    n = 40
    a = rng.uniform(-1, 0, n)
    c = rng.uniform(-1, 0, n)
    bdiag = 2.5 + rng.uniform(0, 1, n)
    d = rng.standard_normal(n)
    A = np.diag(bdiag) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    x = oracle.thomas(a, bdiag, c, d)
    assert np.allclose(x, np.linalg.solve(A, d), rtol=1e-12, atol=1e-13)


def _random_spd_7pt(nx, ny, nz, seed, vscale=1.0):
    """SPD 7-point DIA matrix on an nx x ny x nz grid: a Poisson operator with per-row positive
    scaling, vertical couplings times vscale and a positive diagonal shift."""
    rng = np.random.default_rng(seed)
    off, val = oracle.poisson7pt(nx, ny, nz)
    n = val.shape[1]
    A = oracle.dia_to_dense(off, val)
    nxy = nx * ny
    for r in range(n):
        for c in (r - nxy, r + nxy):
            if 0 <= c < n:
                A[r, c] *= vscale
    A = A - np.diag(np.diag(A))
    A += np.diag(-A.sum(axis=1))                 # weakly dominant, like the flow operator
    w = rng.uniform(0.2, 3.0, n)
    A = A * np.sqrt(np.outer(w, w))
    A += np.diag(rng.uniform(0.01, 1.0, n))
    dense_val = np.zeros_like(val)
    for d, o in enumerate(off):
        for r in range(n):
            c = r + o
            if 0 <= c < n:
                dense_val[d, r] = A[r, c]
    return off, dense_val, A


def test_line_solve_matches_dense_column_systems():
    nx, ny, nz = 4, 3, 7
    off, val, A = _random_spd_7pt(nx, ny, nz, seed=3)
    n, nxy = nx * ny * nz, nx * ny
    T = np.zeros_like(A)                          # the +-nxy diagonals and the main diagonal only
    for r in range(n):
        for c in (r - nxy, r, r + nxy):
            if 0 <= c < n:
                T[r, c] = A[r, c]
    r = np.random.default_rng(4).standard_normal(n)
    z = oracle.line_solve((nx, ny, nz), val, r)
    np.testing.assert_allclose(z, np.linalg.solve(T, r), rtol=1e-12, atol=1e-12 * np.max(np.abs(z)))
    # columns are independent: a residual in one column leaves the others at zero
    r1 = np.zeros(n)
    r1[np.arange(nz) * nxy + 5] = 1.0
    z1 = oracle.line_solve((nx, ny, nz), val, r1)
    assert np.all(z1[np.setdiff1d(np.arange(n), np.arange(nz) * nxy + 5)] == 0.0)


def test_pcg_line_matches_dense_lu_and_beats_jacobi_on_anisotropy():
    nx, ny, nz = 5, 4, 9
    off, val, A = _random_spd_7pt(nx, ny, nz, seed=5, vscale=400.0)   # dz = dx / 20
    n = nx * ny * nz
    b = np.random.default_rng(6).standard_normal(n)
    x_lu = oracle.dense_solve(A, b)
    x, it, rn, conv = oracle.dia_pcg_line((nx, ny, nz), val, b, np.zeros(n), 1000, 1e-13)
    assert conv
    assert np.max(np.abs(x - x_lu)) < 1e-10 * np.max(np.abs(x_lu))
    _, it_j, _, conv_j = oracle.dia_pcg_jacobi(off, val, b, np.zeros(n), 1000, 1e-13)
    assert conv_j and it < it_j / 2, (it, it_j)
