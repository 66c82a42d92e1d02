"""CPU check of the far-field certificate's mathematics (k_cluster.hip): for a
triangle cluster with centre c, radius r, area vector N, moment
M = sum_t (g_t - c) s_t^T and area A, and a point p with d = |c - p| > r,

    | sum_t Omega_t(p) - [N.x + tr M - 3 x^T M x / |x|^2] / |x|^3 | <= 3 r^2 A / (d - r)^4,
    | sum_t Omega_t(p) - N.x / |x|^3 | <= 2 r A / (d - r)^3,     x = c - p,

with Omega_t the Van Oosterom-Strackee solid angle in binary64.  Random
clusters (not necessarily surfaces) and points at 2-24 cluster radii."""
import numpy as np


def _omega(p, a, b, c):
    a, b, c = a - p, b - p, c - p
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    det = a @ np.cross(b, c)
    den = la * lb * lc + (a @ b) * lc + (b @ c) * la + (c @ a) * lb
    return 2 * np.arctan2(det, den)


def test_multipole_bounds_hold():
    rng = np.random.default_rng(1)
    worst2 = worst1 = 0.0
    for _ in range(200):
        n = int(rng.integers(1, 30))
        base = rng.normal(size=3) * 5
        tris = []
        for _ in range(n):
            v0 = base + rng.normal(size=3) * 0.8
            tris.append((v0, v0 + rng.normal(size=3) * 0.5, v0 + rng.normal(size=3) * 0.5))
        V = np.array([v for t in tris for v in t])
        c = 0.5 * (V.min(0) + V.max(0))
        r = np.sqrt(((V - c) ** 2).sum(1).max())
        N = np.zeros(3)
        M = np.zeros((3, 3))
        A = 0.0
        for a, b, cc in tris:
            s = 0.5 * np.cross(b - a, cc - a)
            N += s
            A += np.linalg.norm(s)
            M += np.outer((a + b + cc) / 3 - c, s)
        for beta in (2.0, 3.0, 4.0, 8.0):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            d = r * beta * (1 + rng.random() * 2)
            p = c + u * d
            x = c - p
            xh = x / d
            exact = sum(_omega(p, a, b, cc) for a, b, cc in tris)
            two = (N @ x + np.trace(M) - 3 * xh @ M @ xh) / d ** 3
            one = (N @ x) / d ** 3
            e2 = abs(exact - two) / (3 * r * r * A / (d - r) ** 4)
            e1 = abs(exact - one) / (2 * r * A / (d - r) ** 3)
            worst2, worst1 = max(worst2, e2), max(worst1, e1)
    assert worst2 <= 1.0 and worst1 <= 1.0
    # the bounds are not vacuous: errors come within a factor of 10 of them
    assert worst2 > 0.1 and worst1 > 0.1


def _cluster(rng, n=8):
    base = rng.normal(size=3)
    tris = [(base + rng.normal(size=3), base + rng.normal(size=3), base + rng.normal(size=3)) for _ in range(n)]
    V = np.array([v for t in tris for v in t])
    c = 0.5 * (V.min(0) + V.max(0))
    r = np.sqrt(((V - c) ** 2).sum(1).max())
    N = np.zeros(3)
    M = np.zeros((3, 3))
    for a, b, cc in tris:
        s = 0.5 * np.cross(b - a, cc - a)
        N += s
        M += np.outer((a + b + cc) / 3 - c, s)
    return tris, c, r, N, M


def test_second_term_improves_and_a_planted_sign_error_does_not():
    """Far away the two-term expansion beats the dipole alone; with the second
    term's sign flipped it is worse than the dipole alone."""
    rng = np.random.default_rng(2)
    e_one, e_two, e_bad = [], [], []
    for _ in range(200):
        tris, c, r, N, M = _cluster(rng)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        p = c + u * 8 * r
        x = c - p
        d = np.linalg.norm(x)
        xh = x / d
        exact = sum(_omega(p, a, b, cc) for a, b, cc in tris)
        q = (np.trace(M) - 3 * xh @ M @ xh) / d ** 3
        one = (N @ x) / d ** 3
        e_one.append(abs(exact - one))
        e_two.append(abs(exact - (one + q)))
        e_bad.append(abs(exact - (one - q)))
    assert np.median(e_two) < 0.3 * np.median(e_one)
    assert np.median(e_bad) > np.median(e_one)


def _props(tris, c):
    N = np.zeros(3)
    M = np.zeros((3, 3))
    A = 0.0
    for a, b, cc in tris:
        s = 0.5 * np.cross(b - a, cc - a)
        N += s
        A += np.linalg.norm(s)
        M += np.outer((a + b + cc) / 3 - c, s)
    return N, M, A


def test_parent_cluster_shift_and_bound():
    """A parent made of child clusters (the k_cluster.hip parent build):
    moving each child's moment to the parent centre, M' = sum M_k +
    (c_k - C) N_k^T, equals the moment computed from the triangles about C,
    and the radius max |c_k - C| + r_k keeps the two-term bound valid.  A
    planted transposition of the shift term breaks the identity."""
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(100):
        children = []
        for _ in range(int(rng.integers(2, 9))):
            tris, c, r, N, M = _cluster(rng, n=int(rng.integers(1, 6)))
            off = rng.normal(size=3) * 2
            tris = [(a + off, b + off, cc + off) for a, b, cc in tris]
            children.append((tris, c + off, r, N, M))
        lo = np.min([c - r for _, c, r, _, _ in children], axis=0)
        hi = np.max([c + r for _, c, r, _, _ in children], axis=0)
        C = 0.5 * (lo + hi)
        R = max(np.linalg.norm(c - C) + r for _, c, r, _, _ in children)
        Np = sum(N for _, _, _, N, _ in children)
        Mp = sum(M + np.outer(c - C, N) for _, c, _, N, M in children)
        bad = sum(M + np.outer(N, c - C) for _, c, _, N, M in children)
        all_tris = [t for ch in children for t in ch[0]]
        N2, M2, A = _props(all_tris, C)
        assert np.allclose(Np, N2, atol=1e-9) and np.allclose(Mp, M2, atol=1e-9)
        assert not np.allclose(bad, M2, atol=1e-6)
        V = np.array([v for t in all_tris for v in t])
        assert np.sqrt(((V - C) ** 2).sum(1).max()) <= R + 1e-12
        for beta in (2.0, 4.0):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            d = R * beta * (1 + rng.random())
            p = C + u * d
            x = C - p
            xh = x / d
            exact = sum(_omega(p, a, b, cc) for a, b, cc in all_tris)
            two = (Np @ x + np.trace(Mp) - 3 * xh @ Mp @ xh) / d ** 3
            worst = max(worst, abs(exact - two) / (3 * R * R * A / (d - R) ** 4))
    assert worst <= 1.0
