"""Pins the FP64 oracle against closed forms, invariants and brute force.

The reference ships no golden vectors or known-answer tests for this path,
so every oracle function is checked here against what the mathematics
fixes: closed-form element matrices, exact laminate solutions, symmetry
and invariance of the homogenized tensor, finite differences for the
sensitivities, dense direct solves, adjoint identities and small
pure-Python loops of the published update rules.
"""
import itertools
import math

import numpy as np
import pytest


def _popc(a):
    return bin(a).count("1")


# ---------------------------------------------------------------- template
@pytest.mark.parametrize("n", [4, 16, 128])
def test_template_closed_forms(orc, n):
    Ke, fe, disp = orc.template(n)
    h = 1.0 / n
    P = np.array([[4.0 if a == b else (0.0 if _popc(a ^ b) == 1 else -1.0) for b in range(8)] for a in range(8)])
    np.testing.assert_allclose(Ke, h / 12.0 * P, rtol=0, atol=1e-15 * h)
    # load case k is a unit gradient along grid axis (y, -x, z)
    axis, sign = [1, 0, 2], [1.0, -1.0, 1.0]
    for k in range(3):
        for l in range(8):
            bit = (l >> axis[k]) & 1
            assert fe[l, k] == pytest.approx(sign[k] * (1 if bit else -1) * h * h / 4, abs=1e-15)
            assert disp[l, k] == pytest.approx(sign[k] * bit * h, abs=1e-14)
    np.testing.assert_allclose(Ke @ disp, fe, atol=1e-15)
    w = np.linalg.eigvalsh(Ke)
    assert abs(w[0]) < 1e-14 and w[1] > 1e-3 * h


# ---------------------------------------------------------------- assembly
def _node_index(n, x, y, z):
    return (x % n) + n * ((y % n) + n * (z % n))


def test_assembly_uniform_stencil(orc):
    n = 6
    A = orc.assemble(np.ones((n, n, n)))
    h = 1.0 / n
    for o in range(27):
        d = (o % 3 - 1, o // 3 % 3 - 1, o // 9 - 1)
        nz = sum(1 for v in d if v != 0)
        expect = {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[nz]
        np.testing.assert_allclose(A[o], expect, atol=1e-15)


def test_assembly_symmetry_nullspace_energy(orc):
    rng = np.random.default_rng(1)
    n = 5
    k = rng.uniform(0.01, 1.0, (n, n, n))
    A = orc.assemble(k)
    N = n ** 3
    # symmetry: A[o][v] == A[26-o][v + off(o)]
    for o in range(27):
        dx, dy, dz = o % 3 - 1, o // 3 % 3 - 1, o // 9 - 1
        for v in range(0, N, 7):
            x, y, zz = v % n, v // n % n, v // (n * n)
            w = _node_index(n, x + dx, y + dy, zz + dz)
            assert A[o, v] == pytest.approx(A[26 - o, w], rel=1e-13, abs=1e-16)
    assert np.abs(orc.matvec(k.shape, A, np.ones(N))).max() < 1e-14
    # energy u^T A u equals the element-by-element sum with the pinned Ke
    u = rng.standard_normal((n, n, n))
    Ke, _, _ = orc.template(n)
    e = 0.0
    for l, m in itertools.product(range(8), range(8)):
        ul = np.roll(u, shift=(-(l >> 2 & 1), -(l >> 1 & 1), -(l & 1)), axis=(0, 1, 2))
        um = np.roll(u, shift=(-(m >> 2 & 1), -(m >> 1 & 1), -(m & 1)), axis=(0, 1, 2))
        e += Ke[l, m] * np.sum(k * ul * um)
    uAu = float(u.reshape(-1) @ orc.matvec(k.shape, A, u))
    assert uAu == pytest.approx(e, rel=1e-12)


def test_rhs_closed_forms(orc):
    n = 8
    h = 1.0 / n
    assert np.abs(orc.rhs(np.full((n, n, n), 0.7), 1)).max() < 1e-15
    rng = np.random.default_rng(2)
    kx = rng.uniform(0.1, 1.0, n)
    k = np.broadcast_to(kx[None, None, :], (n, n, n)).copy()
    f = orc.rhs(k, 1).reshape(n, n, n)  # load case 1: gradient along -x
    # f_v = h^2 (kappa(x) - kappa(x-1)) for a laminate in x (4 elements per side)
    expect = h * h * (kx - np.roll(kx, 1))
    np.testing.assert_allclose(f, np.broadcast_to(expect[None, None, :], f.shape), atol=1e-15)
    k3 = rng.uniform(0.1, 1.0, (n, n, n))
    for c in range(3):
        assert abs(orc.rhs(k3, c).sum()) < 1e-14


# ---------------------------------------------------------------- solves
def _laminate(n, axis, vals):
    shape = [1, 1, 1]
    shape[2 - axis] = n  # numpy (z, y, x)
    return np.broadcast_to(np.asarray(vals).reshape(shape), (n, n, n)).copy()


@pytest.mark.parametrize("axis,label", [(0, 1), (1, 0), (2, 2)])
def test_laminate_exact(orc, axis, label):
    n = 8
    vals = np.array([1.0, 1.0, 0.05, 0.05, 0.4, 0.4, 0.9, 0.2])
    C, _, _, _ = orc.homogenize(_laminate(n, axis, vals))
    ve = vals + 1e-9
    harm, arith = 1.0 / np.mean(1.0 / ve), np.mean(ve)
    for i in range(3):
        assert C[i, i] == pytest.approx(harm if i == label else arith, rel=1e-10)
    off = C - np.diag(np.diag(C))
    assert np.abs(off).max() < 1e-12


def test_uniform_and_bounds(orc):
    n = 8
    C, _, _, _ = orc.homogenize(np.full((n, n, n), 0.37))
    np.testing.assert_allclose(C, (0.37 + 1e-9) * np.eye(3), atol=1e-12)
    rng = np.random.default_rng(3)
    k = rng.uniform(1e-3, 1.0, (n, n, n))
    C, _, _, _ = orc.homogenize(k)
    np.testing.assert_allclose(C, C.T, atol=1e-14)
    w = np.linalg.eigvalsh(C)
    ke = k + 1e-9
    assert w[0] >= 1.0 / np.mean(1.0 / ke) - 1e-12
    assert w[-1] <= np.mean(ke) + 1e-12


def test_invariances(orc):
    n = 8
    rng = np.random.default_rng(4)
    k = rng.uniform(1e-3, 1.0, (n, n, n))
    C, _, _, _ = orc.homogenize(k)
    Cs, _, _, _ = orc.homogenize(np.roll(k, (3, -2, 5), axis=(0, 1, 2)))
    np.testing.assert_allclose(Cs, C, atol=1e-12)
    # swap grid x and y: labels 0 <-> 1, C01 kept, C02 -> -C12, C12 -> -C02
    Cp, _, _, _ = orc.homogenize(np.ascontiguousarray(np.swapaxes(k, 1, 2)))
    assert Cp[0, 0] == pytest.approx(C[1, 1], rel=1e-11)
    assert Cp[1, 1] == pytest.approx(C[0, 0], rel=1e-11)
    assert Cp[2, 2] == pytest.approx(C[2, 2], rel=1e-11)
    assert Cp[0, 1] == pytest.approx(C[0, 1], abs=1e-12)
    assert Cp[0, 2] == pytest.approx(-C[1, 2], abs=1e-12)
    assert Cp[1, 2] == pytest.approx(-C[0, 2], abs=1e-12)
    # mirror along grid x: labels (0,1) and (1,2) change sign
    Cm, _, _, _ = orc.homogenize(np.ascontiguousarray(k[:, :, ::-1]))
    assert Cm[0, 1] == pytest.approx(-C[0, 1], abs=1e-12)
    assert Cm[1, 2] == pytest.approx(-C[1, 2], abs=1e-12)
    assert Cm[0, 2] == pytest.approx(C[0, 2], abs=1e-12)
    np.testing.assert_allclose(np.diag(Cm), np.diag(C), rtol=1e-11)


def test_pin_vs_meanzero_and_dense(orc):
    n = 4
    rng = np.random.default_rng(5)
    k = rng.uniform(1e-2, 1.0, (n, n, n))
    C1, chi1, _, _ = orc.homogenize(k, pin=True, want_chi=True)
    C2, chi2, _, _ = orc.homogenize(k, pin=False, want_chi=True)
    np.testing.assert_allclose(C1, C2, atol=1e-13)
    np.testing.assert_allclose(chi1, chi2, atol=1e-11)
    # dense least-squares solve of the assembled system
    A = orc.assemble(k + 1e-9)
    N = n ** 3
    D = np.zeros((N, N))
    for o in range(27):
        dx, dy, dz = o % 3 - 1, o // 3 % 3 - 1, o // 9 - 1
        for v in range(N):
            x, y, z = v % n, v // n % n, v // (n * n)
            D[v, _node_index(n, x + dx, y + dy, z + dz)] += A[o, v]
    for c in range(3):
        f = orc.rhs(k + 1e-9, c)
        sol = np.linalg.lstsq(D, f, rcond=None)[0]
        sol -= sol.mean()
        np.testing.assert_allclose(chi1[c], sol, atol=1e-10)


def test_sensitivity_finite_difference(orc):
    n = 4
    rng = np.random.default_rng(6)
    k = rng.uniform(0.05, 1.0, (n, n, n))
    C, chi, S, _ = orc.homogenize(k, rtol=1e-14, want_chi=True, want_S=True)
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for e in [0, 17, 42, 63]:
        eps = 1e-6
        kp, km = k.copy().reshape(-1), k.copy().reshape(-1)
        kp[e] += eps
        km[e] -= eps
        Cp, _, _, _ = orc.homogenize(kp.reshape(k.shape), rtol=1e-14)
        Cm, _, _, _ = orc.homogenize(km.reshape(k.shape), rtol=1e-14)
        dC = (Cp - Cm) / (2 * eps)
        for c, (i, j) in enumerate(idx):
            assert S[c, e] == pytest.approx(dC[i, j], rel=1e-6, abs=1e-10)
    # Euler identity: sum_e kappa_e S_e = C (the energy is 1-homogeneous in kappa)
    for c, (i, j) in enumerate(idx):
        assert float(np.sum((k.reshape(-1) + 1e-9) * S[c])) == pytest.approx(C[i, j], rel=1e-12, abs=1e-14)
    # contracted sensitivity
    dC = rng.standard_normal((3, 3))
    s = orc.sensitivity(S, dC)
    expect = sum(S[c] * (dC[i, j] + (dC[j, i] if i != j else 0.0)) for c, (i, j) in enumerate(idx))
    np.testing.assert_allclose(s, expect, rtol=1e-13)


# ---------------------------------------------------------------- filters
def test_spline4_taps_and_constants(orc):
    n = 9
    delta = np.zeros((n, n, n))
    delta[4, 4, 4] = 1.0
    out = orc.filter_fwd(delta, kind=0, r=1.5, periodic=False)
    assert np.count_nonzero(out > 0) == 19  # |d|^2 in {0,1,2}
    # Spline4 at q = |d|^2 / r^2 = 4/9 equals 1/9 of the centre weight
    assert out[4, 4, 5] / out[4, 4, 4] == pytest.approx(1.0 / 9.0, rel=1e-12)
    assert out.sum() == pytest.approx(1.0, rel=1e-12)
    const = np.full((5, 6, 7), 0.42)
    np.testing.assert_allclose(orc.filter_fwd(const, 0, 1.5, False), 0.42, rtol=1e-13)
    np.testing.assert_allclose(orc.filter_fwd(const, 1, 2.0, True), 0.42, rtol=1e-13)
    lin = orc.filter_fwd(delta, kind=1, r=2.0, periodic=True)
    assert np.count_nonzero(lin > 0) == 27  # |d|^2 <= 3


@pytest.mark.parametrize("kind,r,periodic", [(0, 1.5, False), (1, 2.0, True), (0, 2.5, False)])
def test_filter_adjoint_identity(orc, kind, r, periodic):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((6, 7, 8))
    y = rng.standard_normal((6, 7, 8))
    lhs = np.sum(orc.filter_fwd(x, kind, r, periodic) * y)
    rhs = np.sum(x * orc.filter_adj(y, kind, r, periodic))
    assert lhs == pytest.approx(rhs, rel=1e-12)


# ---------------------------------------------------------------- OC
def _oc_python(sens, rho, vol, min_rho, step, damp, steps, tol):
    """Pure-Python loop of OCOptimizer::update (OCOptimizer.cu:12-67)."""
    mx = max(abs(s) for s in sens)
    lo, hi = 0.0, mx
    new = list(rho)
    for _ in range(steps):
        g = 0.5 * (lo + hi)
        for i, (s, r) in enumerate(zip(sens, rho)):
            B = -s / g
            if B < 0:
                B = 0.01
            v = B ** damp * r
            v = min(max(v, r - step), r + step)
            new[i] = min(max(v, min_rho), 1.0)
        cur = sum(new) / len(new)
        if cur < vol - tol:
            hi = g
        elif cur > vol + tol:
            lo = g
        else:
            break
    return new


def test_oc_update_matches_loop(orc):
    rng = np.random.default_rng(8)
    n = 64
    rho = rng.uniform(0.05, 0.9, n)
    sens = -rng.uniform(0.0, 2.0, n)
    out, steps, g = orc.oc_update(sens, rho, 0.4, 0.001, 0.2, 0.5, 30, 1e-6)
    np.testing.assert_allclose(out, _oc_python(sens, rho, 0.4, 0.001, 0.2, 0.5, 30, 1e-6), rtol=1e-14)
    assert np.all(np.abs(out - rho) <= 0.2 + 1e-15)
    assert abs(out.mean() - 0.4) < 1e-3


def test_oc_filter_sens_loop(orc):
    rng = np.random.default_rng(9)
    n = 4
    rho = rng.uniform(0.1, 1.0, (n, n, n))
    sens = rng.standard_normal((n, n, n))
    out = orc.oc_filter_sens(rho, sens, 2.0)
    ref = np.zeros_like(rho)
    for z, y, x in itertools.product(range(n), repeat=3):
        s = ws = 0.0
        for dz, dy, dx in itertools.product(range(-2, 3), repeat=3):
            d = math.sqrt(dx * dx + dy * dy + dz * dz)
            w = max(0.0, 1.0 - d / 2.0)
            Z, Y, X = (z + dz) % n, (y + dy) % n, (x + dx) % n
            s += w * rho[Z, Y, X] * sens[Z, Y, X]
            ws += w
        ref[z, y, x] = s / (ws * rho[z, y, x])
    np.testing.assert_allclose(out, ref, rtol=1e-12)


# ---------------------------------------------------------------- MMA
def _mma_subproblem(x, xmin, xmax, df0, g, dg, move=1.0):
    """Approximation built at iteration 0 (asymptotes x -+ 0.5 (xmax - xmin))."""
    dxm = xmax - xmin
    low, upp = x - 0.5 * dxm, x + 0.5 * dxm
    alfa = np.maximum(np.maximum(low + 0.1 * (x - low), x - move * dxm), xmin)
    beta = np.minimum(np.minimum(upp - 0.1 * (upp - x), x + move * dxm), xmax)
    inv = 1.0 / np.maximum(dxm, 1e-5)
    ux1, xl1 = upp - x, x - low

    def pq(gr):
        p, q = np.maximum(gr, 0), np.maximum(-gr, 0)
        r = 1e-3 * (p + q) + 1e-5 * inv
        return (p + r) * ux1 ** 2, (q + r) * xl1 ** 2
    p0, q0 = pq(df0)
    P, Q = zip(*[pq(dg[i]) for i in range(dg.shape[0])])
    P, Q = np.array(P), np.array(Q)
    b = (P / ux1 + Q / xl1).sum(axis=1) - g
    return low, upp, alfa, beta, p0, q0, P, Q, b


def test_mma_subproblem_against_dual_solver(orc):
    """One oracle MMA update equals the minimiser of the MMA subproblem found
    independently by maximising its concave dual over lambda >= 0."""
    from scipy.optimize import minimize
    rng = np.random.default_rng(21)
    n, m = 300, 2
    c, d = 1e3, 1.0
    x = rng.uniform(0.2, 0.8, n)
    df0 = rng.standard_normal(n)
    dg = rng.standard_normal((m, n)) / n + 1.0 / n
    g = np.array([0.05, -0.02])
    low, upp, alfa, beta, p0, q0, P, Q, b = _mma_subproblem(x, 0.0, 1.0, df0, g, dg)

    def xstar(lam):
        A = p0 + lam @ P
        B = q0 + lam @ Q
        xs = (low * np.sqrt(A) + upp * np.sqrt(B)) / (np.sqrt(A) + np.sqrt(B))
        return np.clip(xs, alfa, beta)

    def negdual(lam):
        lam = np.maximum(lam, 0)
        xs = xstar(lam)
        A = p0 + lam @ P
        B = q0 + lam @ Q
        y = np.maximum(0, (lam - c) / d)
        W = np.sum(A / (upp - xs) + B / (xs - low)) + np.sum(c * y + 0.5 * d * y * y - lam * y) - lam @ b
        grad = (P / (upp - xs) + Q / (xs - low)).sum(axis=1) - y - b
        return -W, -grad
    res = minimize(negdual, np.ones(m), jac=True, method="L-BFGS-B", bounds=[(0, None)] * m,
                   options={"gtol": 1e-13, "ftol": 1e-16, "maxiter": 5000})
    x_dual = xstar(np.maximum(res.x, 0))
    # driven to a tiny barrier the interior point lands on the dual solution
    tight = orc.MMA(m, n, 1.0, 0.0, c, d, 0.0, 1.0)
    tight.epsimin = 1e-13
    x_tight, _ = tight.update(0, x, df0, g, dg)
    np.testing.assert_allclose(x_tight, x_dual, atol=1e-8)
    # at the reference's last barrier level (1e-7) it stops within 1e-4
    mma = orc.MMA(m, n, 1.0, 0.0, c, d, 0.0, 1.0)
    x_new, its = mma.update(0, x, df0, g, dg)
    assert its > 0
    assert np.abs(x_new - x_dual).max() < 1e-4


def test_mma_converges_to_constrained_optimum(orc):
    """min sum (x-t)^2 s.t. mean(x) <= 0.3, mean(w x) <= 0.25 on [0,1]^n: the
    MMA iterates reach the optimum computed by SLSQP."""
    from scipy.optimize import minimize
    rng = np.random.default_rng(22)
    n = 40
    t = rng.uniform(0, 1, n)
    w = rng.uniform(0.5, 1.5, n)
    cons = [{"type": "ineq", "fun": lambda x: 0.3 - x.mean(), "jac": lambda x: -np.ones(n) / n},
            {"type": "ineq", "fun": lambda x: 0.25 - (w * x).mean(), "jac": lambda x: -w / n}]
    ref = minimize(lambda x: ((x - t) ** 2).sum(), np.full(n, 0.2), jac=lambda x: 2 * (x - t),
                   bounds=[(0, 1)] * n, constraints=cons, method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
    mma = orc.MMA(2, n, 1.0, 0.0, 1e3, 1.0, 0.0, 1.0)
    x = np.full(n, 0.2)
    for it in range(60):
        g = np.array([x.mean() - 0.3, (w * x).mean() - 0.25])
        dg = np.stack([np.ones(n) / n, w / n])
        x, _ = mma.update(it, x, 2 * (x - t), g, dg)
    assert np.abs(x - ref.x).max() < 2e-3


def test_symmetrize_group_properties(orc):
    """Orbit averages: group sizes, idempotence, invariance of the result
    under every group element, mean preservation, symmetric input
    unchanged, and the 1-D closed form of a reflection."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal((6, 6, 6))
    sizes = {"reflect3": 8, "reflect6": 48, "rotate3": 24, "rotate2": 2}
    for sym, size in sizes.items():
        G = orc.sym_group(sym)
        assert len(G) == size and len(set(G)) == size
        y = orc.symmetrize(x, sym)
        np.testing.assert_allclose(orc.symmetrize(y, sym), y, atol=1e-12)
        for g in G:
            np.testing.assert_allclose(orc.sym_apply(y, g), y, atol=1e-12)
        assert abs(y.mean() - x.mean()) < 1e-12
    # rotate3 elements are rotations: determinant +1 of the signed permutation
    for perm, flip in orc.sym_group("rotate3"):
        M = np.zeros((3, 3))
        for j in range(3):
            M[j, perm[j]] = -1.0 if (flip >> j) & 1 else 1.0
        assert abs(np.linalg.det(M) - 1.0) < 1e-12
    # a field depending on x only: reflect3 gives (f(x) + f(n-1-x)) / 2
    f = rng.standard_normal(6)
    xf = np.broadcast_to(f[None, None, :], (4, 5, 6))
    np.testing.assert_allclose(orc.symmetrize(xf, "reflect3"), np.broadcast_to(0.5 * (f + f[::-1]), (4, 5, 6)),
                               atol=1e-12)
    # a radially symmetric field about the centre is invariant under all groups
    c = (np.arange(6) - 2.5) ** 2
    r = c[:, None, None] + c[None, :, None] + c[None, None, :]
    for sym in sizes:
        np.testing.assert_allclose(orc.symmetrize(r, sym), r, atol=1e-12)
