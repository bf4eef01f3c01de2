"""Pins for the oracle functions that no reference output fixes:

* the dogleg / 2-D subspace step (PSBA/trust_region.cpp:520-595) against a
  restatement from Nocedal & Wright, Numerical Optimization (2nd ed.), ch. 4:
  the subspace minimiser of m(p) = g.p + p.Bp/2 from its own 2x2 normal
  equations and by direct minimisation, the scaled Cauchy step, the full step
  and the dogleg boundary point (section 4.1, eq. 4.16);
* the modified Cholesky (PSBA/cl_cholmod.cpp:109-202, cholmod_blk.cl) against
  a column-by-column Gill-Murray-Wright restatement with delta and beta
  recomputed from cl_cholmod.cpp:161-164, and the bounds it promises;
* the small-determinant branch of the 3x3 inverse (compute_Vinv.cl:31-73) on
  hand-built matrices whose results were derived by hand (comments).
"""
import numpy as np
import pytest
from scipy.optimize import minimize

# ---------------------------------------------------------------------------
# dogleg step
# ---------------------------------------------------------------------------
def dogleg_instance(rng, N, lam, scale_delta):
    J = rng.standard_normal((N + 5, N))
    J[:, 0] *= rng.uniform(0.05, 1.0)  # some anisotropy
    B = J.T @ J
    g = rng.standard_normal(N)
    P_U = -(g @ g) / (g @ B @ g) * g                      # N&W eq. 4.12 (unbounded Cauchy step)
    P_B = -np.linalg.solve(B + lam * np.eye(N), g)        # full (damped) step
    delta = scale_delta * np.linalg.norm(P_B)
    return B, g, P_U, P_B, delta


def textbook_step(B, g, P_U, P_B, delta):
    """(p, branch, reported norm) from N&W ch. 4 with the reference's branch
    order: the minimiser of m over span{P_U, P_B} if it lies in the region,
    else the scaled Cauchy step when P_U leaves the region, else P_B when it
    lies inside, else the point where the dogleg path P_U -> P_B meets the
    boundary.  The reported norm follows PSBA/trust_region.cpp:564-569."""
    # subspace minimiser: grad of m(a P_U + b P_B) = 0, a 2x2 SPD system
    V = np.stack([P_U, P_B], axis=1)
    eta = np.linalg.solve(V.T @ B @ V, -(V.T @ g))
    p2 = V @ eta
    n2 = np.linalg.norm(p2)
    if n2 <= delta:
        return p2, "subspace", n2
    if np.linalg.norm(P_U) > delta:
        return -delta * g / np.linalg.norm(g), "cauchy", delta
    if np.linalg.norm(P_B) <= delta:
        return P_B, "full", np.sqrt(n2 + P_B @ P_B)
    # |P_U + t (P_B - P_U)| = delta for t in [0, 1] (eq. 4.16 with t = tau - 1)
    d = P_B - P_U
    roots = np.roots([d @ d, 2 * (P_U @ d), P_U @ P_U - delta ** 2])
    t = [r.real for r in roots if abs(r.imag) < 1e-12 and -1e-12 <= r.real <= 1 + 1e-12]
    assert len(t) == 1  # the path crosses the boundary once (N&W lemma 4.2)
    return P_U + t[0] * d, "dogleg", delta


def test_dogleg_step_against_nocedal_wright(O):
    rng = np.random.default_rng(20)
    seen = set()
    for trial in range(400):
        N = int(rng.integers(3, 12))
        lam = float(rng.choice([0.0, 0.0, rng.uniform(0.01, 5.0)]))
        B, g, P_U, P_B, delta = dogleg_instance(rng, N, lam, float(rng.uniform(0.05, 1.6)))
        p, nrm = O.dogleg_step(P_U, P_B, g, P_U @ B @ P_U, P_U @ B @ P_B, P_B @ B @ P_B, delta)
        pe, branch, ne = textbook_step(B, g, P_U, P_B, delta)
        seen.add(branch)
        scale = max(np.linalg.norm(pe), 1e-300)
        assert np.linalg.norm(p - pe) <= 1e-9 * scale, (trial, branch)
        assert abs(nrm - ne) <= 1e-9 * ne, (trial, branch)
        if branch == "subspace":
            # first-order optimality within the subspace
            r = g + B @ p
            assert abs(P_U @ r) <= 1e-8 * np.linalg.norm(P_U) * np.linalg.norm(g)
            assert abs(P_B @ r) <= 1e-8 * np.linalg.norm(P_B) * np.linalg.norm(g)
        elif branch == "full":
            assert np.linalg.norm(p) <= delta
        else:  # scaled Cauchy and dogleg steps end on the boundary
            assert abs(np.linalg.norm(p) - delta) <= 1e-10 * delta, (trial, branch)
    assert seen == {"subspace", "cauchy", "full", "dogleg"}, seen


def test_dogleg_subspace_minimiser_by_direct_minimisation(O):
    """The subspace step minimises m over span{P_U, P_B} (checked with a
    general-purpose minimiser, not the 2x2 formula)."""
    rng = np.random.default_rng(21)
    for _ in range(20):
        B, g, P_U, P_B, _ = dogleg_instance(rng, 8, float(rng.uniform(0.1, 3.0)), 1.0)
        p, _ = O.dogleg_step(P_U, P_B, g, P_U @ B @ P_U, P_U @ B @ P_B, P_B @ B @ P_B, 1e30)

        def m(ab):
            q = ab[0] * P_U + ab[1] * P_B
            return g @ q + 0.5 * q @ B @ q

        res = minimize(m, np.array([0.3, 0.3]), method="BFGS", options=dict(gtol=1e-12))
        q = res.x[0] * P_U + res.x[1] * P_B
        assert np.linalg.norm(p - q) <= 1e-5 * np.linalg.norm(q)
        assert m(np.linalg.lstsq(np.stack([P_U, P_B], 1), p, rcond=None)[0]) <= res.fun + 1e-12 * abs(res.fun)


def test_dogleg_undamped_full_step_inside_region(O):
    """With P_B the exact Newton step the subspace minimiser is P_B itself
    (it minimises m globally and lies in the span)."""
    rng = np.random.default_rng(22)
    B, g, P_U, P_B, _ = dogleg_instance(rng, 10, 0.0, 1.0)
    p, nrm = O.dogleg_step(P_U, P_B, g, P_U @ B @ P_U, P_U @ B @ P_B, P_B @ B @ P_B, 2 * np.linalg.norm(P_B))
    assert np.allclose(p, P_B, rtol=1e-10, atol=1e-12 * np.linalg.norm(P_B))
    assert abs(nrm - np.linalg.norm(P_B)) <= 1e-10 * nrm


# ---------------------------------------------------------------------------
# modified Cholesky
# ---------------------------------------------------------------------------
def delta_beta(A):
    """PSBA/cl_cholmod.cpp:161-164 with xi = max |off-diagonal| and
    gamma = max |diagonal| (the two passes of kern_delta_beta,
    cholmod_blk.cl:797-823)."""
    n = A.shape[0]
    off = np.abs(A - np.diag(np.diag(A)))
    xi, gamma = off.max(), np.abs(np.diag(A)).max()
    delta = 1e-15 * max(xi + gamma, 1.0)
    beta = np.sqrt(max(max(gamma, 1e-15), xi / np.sqrt(n * n - 1)))
    return delta, beta


def gmw_columns(A, delta, beta):
    """Gill, Murray & Wright (1981) modified Cholesky, one column at a time:
    c_jj = a_jj - sum_k l_jk^2, d_j = max(|c_jj|, delta), l_ij = c_ij / sqrt(d_j);
    when some l_ij exceeds beta (the reference's signed test, cholmod_blk.cl:549)
    sqrt(d_j) is replaced by theta_j / beta, theta_j = max_i |c_ij|."""
    n = A.shape[0]
    L = np.zeros_like(A)
    for j in range(n):
        c = A[j:, j] - L[j:, :j] @ L[j, :j]
        d = max(abs(c[0]), delta)
        col = c[1:] / np.sqrt(d)
        sd = np.sqrt(d)
        if col.size and col.max() > beta:
            sd = np.abs(c[1:]).max() / beta
            col = c[1:] / sd
        L[j, j] = sd
        L[j + 1:, j] = col
    E = np.sum(L * L, axis=1) - np.diag(A)
    return L, E


def indefinite(rng, n, neg_diag):
    X = rng.standard_normal((n, n))
    A = X + X.T
    if neg_diag:
        A[np.diag_indices(n)] = -np.abs(A[np.diag_indices(n)]) - 0.5
    return A


@pytest.mark.parametrize("n", [6, 12, 30])
def test_modchol_every_block_fails_matches_gmw(O, n):
    """Negative diagonal entries make every 3x3 block step fail at its first
    pivot, so the whole factor comes from the column path: it must equal the
    GMW restatement with delta and beta recomputed here."""
    rng = np.random.default_rng(30 + n)
    for _ in range(5):
        A = indefinite(rng, n, True)
        lam, L, E = O.dense_modchol(A)
        delta, beta = delta_beta(A)
        Le, Ee = gmw_columns(A, delta, beta)
        assert np.allclose(L, Le, rtol=1e-12, atol=1e-12 * np.abs(Le).max())
        assert np.allclose(E, Ee, rtol=1e-10, atol=1e-10 * np.abs(Ee).max())
        assert abs(lam - abs(E.sum()) / n) <= 1e-12 * lam
        # GMW bounds: every pivot at least delta, every entry below the
        # diagonal at most beta (signed, as the reference tests it)
        assert np.all(np.diag(L) ** 2 >= delta * (1 - 1e-12))
        assert np.all(np.tril(L, -1) <= beta * (1 + 1e-12))


@pytest.mark.parametrize("n", [9, 21, 40])
def test_modchol_bounds_on_mixed_paths(O, n):
    """Indefinite matrices where some blocks take the 3x3 step and some fall
    back: L L^T = A + diag(E) with E >= 0, every strictly lower entry <= beta
    (both paths enforce it), and a modification only where the plain
    factorisation breaks down."""
    rng = np.random.default_rng(40 + n)
    for _ in range(5):
        A = indefinite(rng, n, False) + n * 0.3 * np.eye(n)
        lam, L, E = O.dense_modchol(A)
        delta, beta = delta_beta(A)
        assert np.allclose(L @ L.T, A + np.diag(E), rtol=0, atol=1e-10 * np.abs(A).max())
        assert np.all(E >= -1e-10 * np.abs(A).max())
        assert np.all(np.tril(L, -1) <= beta * (1 + 1e-12))
    # a wrong beta changes the factor: the restatement with beta * 1.01 differs
    A = indefinite(np.random.default_rng(7), 12, True)
    _, L, _ = O.dense_modchol(A)
    delta, beta = delta_beta(A)
    assert not np.allclose(L, gmw_columns(A, delta, beta * 1.01)[0], rtol=1e-9, atol=0)
    assert np.allclose(L, gmw_columns(A, delta, beta)[0], rtol=1e-12, atol=1e-12 * np.abs(L).max())


# ---------------------------------------------------------------------------
# small-determinant branch of the 3x3 inverse
# ---------------------------------------------------------------------------
S = 1e-6  # scale: every case below has |det| ~ 1e-18 < 1e-16, so the branch is taken

# Worked by hand through compute_Vinv.cl:31-73 (signed pivot tests, rows
# swapped in place, D = u11 u22 u33 of the resulting elimination, result
# adj(A) / D with adj from the unswapped entries):
#  diag(1,2,3) s: no swap; D = 6 s^3 = det                     -> +A^-1
#  [[2,1,0],[1,2,0],[0,0,1]] s: no swap (2 > 1, 2 > 0); u = 2, 3/2, 1;
#      D = 3 s^3 = det                                         -> +A^-1
#  [[1,2,0],[2,5,1],[0,1,3]] s: 1 < 2 swaps rows 0,1; l = 1/2, 0; rows 1,2
#      become (-1/2, -1/2) and (1, 3); -1/2 < 1 swaps them; l = -1/2,
#      u33 = -1/2 + 3/2 = 1; D = 2 * 1 * 1 s^3 = det (two swaps) -> +A^-1
#  [[1,2,0],[2,5,-1],[0,-1,2]] s: rows 0,1 swapped; rows 1,2 become
#      (-1/2, 1/2) and (-1, 2); -1/2 < -1 is false, no swap; l = 2,
#      u33 = 2 - 1 = 1; D = 2 (-1/2) 1 s^3 = -det (one swap)  -> -A^-1
#  [[1,2,0],[2,5,0],[0,0,1]] s: rows 0,1 swapped; rows 1,2 become
#      (-1/2, 0) and (0, 1); -1/2 < 0 swaps them, leaving u22 = 0; no
#      elimination; D = 2 * 0 * 0 = 0                          -> adj / 0
HAND = [
    (np.diag([1.0, 2.0, 3.0]), +1),
    (np.array([[2.0, 1, 0], [1, 2, 0], [0, 0, 1]]), +1),
    (np.array([[1.0, 2, 0], [2, 5, 1], [0, 1, 3]]), +1),
    (np.array([[1.0, 2, 0], [2, 5, -1], [0, -1, 2]]), -1),
]


@pytest.mark.parametrize("k", range(len(HAND)))
def test_vinv_small_determinant_branch(O, k):
    M, sign = HAND[k]
    A = S * M
    inv, branch = O.vinv3(A)
    assert branch
    assert np.allclose(inv, sign * np.linalg.inv(A), rtol=1e-12, atol=0)


def test_vinv_branch_zero_pivot_and_damping(O):
    A = S * np.array([[1.0, 2, 0], [2, 5, 0], [0, 0, 1]])
    inv, branch = O.vinv3(A)
    assert branch
    assert not np.all(np.isfinite(inv))  # adj(A) / 0
    # damping enters before the branch test: the same block with mu = 1 is regular
    inv, branch = O.vinv3(A, 1.0)
    assert not branch
    assert np.allclose(inv, np.linalg.inv(A + np.eye(3)), rtol=1e-13, atol=0)
    # a regular block takes the cofactor formula
    B = np.array([[4.0, 1, 0.5], [1, 3, 0.2], [0.5, 0.2, 2]])
    inv, branch = O.vinv3(B)
    assert not branch and np.allclose(inv, np.linalg.inv(B), rtol=1e-14, atol=0)
