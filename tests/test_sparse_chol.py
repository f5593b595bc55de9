"""The host sparse LDL^T (precond/sparse_chol.rs) through the C ABI against dense linear algebra:
exact solves of SPD systems, the grounded generalized inverse of singular Laplacians (one
grounded pivot per connected component, A M A = A), and the reduced-system matrices of FE
designs assembled from dummies (operator.rs:2292-2315's explicit S)."""
import numpy as np
import pytest

from fereg_amd.sparse_chol import SparseChol


def _laplacian(n, edges, w):
    a = np.zeros((n, n))
    for (i, j), x in zip(edges, w):
        a[i, i] += x
        a[j, j] += x
        a[i, j] -= x
        a[j, i] -= x
    return a


def test_spd_solve_matches_dense():
    rng = np.random.default_rng(0)
    n = 60
    b = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.08)
    a = b @ b.T + n * np.eye(n)
    f = SparseChol.from_dense(a)
    assert f.grounded == 0
    rhs = rng.standard_normal((n, 3))
    np.testing.assert_allclose(f.solve(rhs), np.linalg.solve(a, rhs), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(f.solve(rhs[:, 0]), np.linalg.solve(a, rhs[:, 0]), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nlast", [0, 5])
def test_singular_laplacian_is_grounded_once_per_component(nlast):
    # three components: a ring, a path and a random graph; the solve is a reflexive generalized
    # inverse (A M A = A) and solves every consistent system exactly
    rng = np.random.default_rng(1)
    edges = [(i, (i + 1) % 40) for i in range(40)]
    edges += [(40 + i, 41 + i) for i in range(24)]
    for _ in range(80):
        i, j = rng.choice(np.arange(65, 95), 2, replace=False)
        edges.append((i, j))
    n = 95
    a = _laplacian(n, edges, rng.uniform(0.5, 2.0, len(edges)))
    f = SparseChol.from_dense(a, nlast=nlast)
    assert f.grounded == 3
    m = f.solve(np.eye(n))
    np.testing.assert_allclose(a @ m @ a, a, atol=1e-9 * np.abs(a).max())
    v = rng.standard_normal((n, 2))
    x = f.solve(a @ v)
    np.testing.assert_allclose(a @ x, a @ v, atol=1e-9 * np.abs(a @ v).max())


def test_reduced_fe_matrix_from_dummies():
    # S = A_rr - A_r1 A_11^-1 A_1r for an (indiv, firm, year) design: the Schur complement of the
    # eliminated FE1, with small coordinates (years) ordered last
    rng = np.random.default_rng(2)
    n, g1, F, T = 3000, 300, 120, 6
    ind = rng.integers(0, g1, n)
    firm = rng.integers(0, F, n)
    yr = rng.integers(0, T, n)
    w = rng.uniform(0.5, 2.0, n)
    D1 = np.zeros((n, g1)); D1[np.arange(n), ind] = 1
    Dr = np.zeros((n, T + F)); Dr[np.arange(n), yr] = 1; Dr[np.arange(n), T + firm] = 1
    n1 = D1.T @ w
    Arr = Dr.T @ (w[:, None] * Dr)
    Ar1 = Dr.T @ (w[:, None] * D1)
    S = Arr - (Ar1 / n1) @ Ar1.T
    f = SparseChol.from_dense(S, nlast=T)
    # year and firm each carry a constant that FE1 absorbs: two null directions
    assert f.grounded == S.shape[0] - np.linalg.matrix_rank(S) == 2
    v = rng.standard_normal((T + F, 2))
    x = f.solve(S @ v)
    np.testing.assert_allclose(S @ x, S @ v, atol=1e-9 * np.abs(S @ v).max())


def test_rejects_entries_above_the_diagonal():
    import fereg_amd as F
    with pytest.raises(F.FeregError):
        SparseChol(2, [0, 2, 3], [0, 1, 1], [1.0, 0.5, 1.0])


def _sparse_laplacian(n, edges, w):
    import scipy.sparse as sp
    i = np.array([e[0] for e in edges]); j = np.array([e[1] for e in edges])
    a = sp.coo_matrix((np.concatenate([w, w, -w, -w]),
                       (np.concatenate([i, j, i, j]), np.concatenate([i, j, j, i]))), shape=(n, n)).tocsr()
    a.sum_duplicates()
    return a


def _from_csr(a, nlast=0):
    lo = __import__("scipy.sparse", fromlist=["tril"]).tril(a).tocsr()
    lo.sort_indices()
    return SparseChol(a.shape[0], lo.indptr.astype(np.uint64), lo.indices, lo.data, nlast)


def test_long_ring_is_dissected():
    """Minimum degree leaves a chain as deep as the ring; nested dissection cuts it to about
    log2(n) separators, and the solve is unchanged."""
    rng = np.random.default_rng(3)
    n = 6000
    edges = [(i, (i + 1) % n) for i in range(n)]
    a = _sparse_laplacian(n, edges, rng.uniform(0.5, 2.0, n))
    f = _from_csr(a)
    assert f.grounded == 1
    assert f.height < 400, f.height
    v = rng.standard_normal((n, 2))
    x = f.solve(a @ v)
    np.testing.assert_allclose(a @ x, a @ v, atol=1e-9 * np.abs(a @ v).max())


def test_disconnected_rings_with_rows_ordered_last():
    rng = np.random.default_rng(4)
    n = 3000
    edges = [(i, (i + 1) % 1000) for i in range(1000)] + [(1000 + i, 1000 + (i + 1) % 2000) for i in range(2000)]
    hub = [(k, rng.integers(0, n)) for k in range(3) for _ in range(50)]  # rows 0..2: dense hubs
    edges += [(int(i), int(j)) for i, j in hub if i != j]
    a = _sparse_laplacian(n, edges, rng.uniform(0.5, 2.0, len(edges)))
    f = _from_csr(a, nlast=3)
    assert f.grounded == n - np.linalg.matrix_rank(a.toarray())
    v = rng.standard_normal((n, 3))
    x = f.solve(a @ v)
    np.testing.assert_allclose(a @ x, a @ v, atol=1e-9 * np.abs(a @ v).max())


def test_selected_inverse_matches_the_dense_inverse_on_the_pattern():
    """Takahashi's recurrences give the inverse's entries on every structural pair (and the
    diagonal) of an SPD matrix; for a grounded Laplacian, u'Z u on range vectors is invariant."""
    rng = np.random.default_rng(5)
    n = 80
    b = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.06)
    a = b @ b.T + 0.5 * np.eye(n)
    f = SparseChol.from_dense(a)
    inv = np.linalg.inv(a)
    r, c = np.nonzero(np.tril(a))
    np.testing.assert_allclose(f.inverse_entries(r, c), inv[r, c], rtol=1e-10, atol=1e-13)
    # a connected ring (one null direction): for u = e_i - e_j with i, j adjacent (a range
    # vector), u'Zu = u'A^+u
    m = 200
    edges = [(i, (i + 1) % m) for i in range(m)]
    L = _sparse_laplacian(m, edges, rng.uniform(0.5, 2.0, m)).toarray()
    g = SparseChol.from_dense(L)
    P = np.linalg.pinv(L)
    for i in range(0, m, 17):
        j = (i + 1) % m
        z = g.inverse_entries([i, j, i], [i, j, j])
        assert abs((z[0] + z[1] - 2 * z[2]) - (P[i, i] + P[j, j] - 2 * P[i, j])) < 1e-9 * abs(P[i, i])
