"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. Arrays are float64 with x
fastest, i.e. numpy shape (nz, ny, nx).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle/liboracle.so missing: run make -C oracle")
        L = ctypes.CDLL(path)
        i, d, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.orc_template.argtypes = [d, _dp, _dp, _dp]
        L.orc_assemble.argtypes = [i, i, i, _dp, _dp, _dp]
        L.orc_matvec.argtypes = [i, i, i, _dp, _dp, _dp]
        L.orc_rhs.argtypes = [i, i, i, _dp, _dp, i, _dp]
        L.orc_cg.argtypes = [i, i, i, _dp, _dp, _dp, d, i, i, ctypes.POINTER(d)]
        L.orc_cg.restype = i
        L.orc_homogenize.argtypes = [i, i, i, _dp, d, i, vp, _dp, vp]
        L.orc_homogenize.restype = i
        L.orc_element_energy.argtypes = [i, i, i, _dp, _dp]
        L.orc_sensitivity.argtypes = [i, _dp, _dp, _dp]
        L.orc_filter_fwd.argtypes = [i, i, i, i, d, i, _dp, _dp, vp]
        L.orc_filter_adj.argtypes = [i, i, i, i, d, i, _dp, _dp]
        L.orc_oc_filter_sens.argtypes = [i, i, i, d, _dp, _dp, _dp]
        L.orc_oc_update.argtypes = [i, _dp, _dp, _dp, d, d, d, d, i, d, ctypes.POINTER(d)]
        L.orc_oc_update.restype = i
        L.orc_mmasub.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, d, _dp, _dp, _dp, d, d]
        L.orc_mmasub.restype = i
        L.orc_mma_stats.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
        L.orc_mma_stats.restype = None
        L.orc_mma_stats_reset.argtypes = []
        L.orc_mma_stats_reset.restype = None
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_gs_ell.argtypes = [i, i, i, _dp, _dp, _dp, i, i, i]
        L.orc_brick_gs_ell.argtypes = [i, i, i, _dp, _dp, _dp, _dp, i, i, i, i, i, i]
        L.orc_galerkin_ell.argtypes = [i, i, i, _dp, _dp]
        for nm in ("orc_mg_homogenize", "orc32_mg_homogenize"):
            f = getattr(L, nm)
            f.argtypes = [i, i, i, _dp, d, vp, _dp, vp, ip, _dp]
            f.restype = i
        _LIB = L
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def template(n):
    Ke, fe, disp = np.zeros(64), np.zeros(24), np.zeros(24)
    lib().orc_template(0.5 / n, Ke, fe, disp)
    return Ke.reshape(8, 8), fe.reshape(8, 3), disp.reshape(8, 3)


def assemble(kappa_eff):
    k = _d(kappa_eff)
    nz, ny, nx = k.shape
    Ke, _, _ = template(nx)
    A = np.zeros(27 * k.size)
    lib().orc_assemble(nx, ny, nz, k.reshape(-1), _d(Ke).reshape(-1), A)
    return A.reshape(27, k.size)


def matvec(shape, A, x):
    nz, ny, nx = shape
    y = np.zeros(nx * ny * nz)
    lib().orc_matvec(nx, ny, nz, _d(A).reshape(-1), _d(x).reshape(-1), y)
    return y


def rhs(kappa_eff, k):
    ke = _d(kappa_eff)
    nz, ny, nx = ke.shape
    _, fe, _ = template(nx)
    f = np.zeros(ke.size)
    lib().orc_rhs(nx, ny, nz, ke.reshape(-1), _d(fe).reshape(-1), k, f)
    return f


def cg(shape, A, b, rtol=1e-12, pin=True, maxit=20000):
    """FP64 Jacobi-PCG on the ELL-27 operator; returns (x, iterations, relres)."""
    nz, ny, nx = shape
    x = np.zeros(nx * ny * nz)
    rr = ctypes.c_double()
    it = lib().orc_cg(nx, ny, nz, _d(A).reshape(-1), _d(b).reshape(-1), x, rtol, maxit, 1 if pin else 0,
                      ctypes.byref(rr))
    return x, it, rr.value


def homogenize(kappa, rtol=1e-12, pin=True, want_chi=False, want_S=False):
    """FP64 kappa^H of the SIMP-mapped conductivity `kappa` (1e-9 floor added
    inside). Returns (C, chi or None, S or None, iterations)."""
    k = _d(kappa)
    nz, ny, nx = k.shape
    n = k.size
    C = np.zeros(9)
    chi = np.zeros(3 * n) if want_chi else None
    S = np.zeros(6 * n) if want_S else None
    cp = chi.ctypes.data_as(ctypes.c_void_p) if want_chi else None
    sp = S.ctypes.data_as(ctypes.c_void_p) if want_S else None
    it = lib().orc_homogenize(nx, ny, nz, k.reshape(-1), rtol, 1 if pin else 0, cp, C, sp)
    return (C.reshape(3, 3), None if chi is None else chi.reshape(3, n),
            None if S is None else S.reshape(6, n), it)


def gs(shape, A, f, u, sweeps=1, reverse=False):
    """Exact multicolour Gauss-Seidel in the reference's colour order (reverse:
    7..0) on the ELL-27 operator A of a lattice of `shape` (nz, ny, nx); f and
    u hold nrhs right-hand sides (nrhs, n). Returns the new u."""
    nz, ny, nx = shape
    fv = _d(f).reshape(-1, nx * ny * nz)
    out = _d(u).reshape(fv.shape).copy()
    lib().orc_gs_ell(nx, ny, nz, _d(A).reshape(-1), fv.reshape(-1), out.reshape(-1), int(sweeps),
                     1 if reverse else 0, fv.shape[0])
    return out


def brick_gs(shape, A, f, u, brick, sweeps=1, reverse=False):
    """Brick Gauss-Seidel (bricks of brick = (bx, by, bz) nodes, couplings
    leaving a brick frozen at the entry u); returns the new u."""
    nz, ny, nx = shape
    fv = _d(f).reshape(-1, nx * ny * nz)
    uin = _d(u).reshape(fv.shape)
    out = np.zeros_like(uin)
    bx, by, bz = brick
    lib().orc_brick_gs_ell(nx, ny, nz, _d(A).reshape(-1), fv.reshape(-1), uin.reshape(-1), out.reshape(-1),
                           bx, by, bz, int(sweeps), 1 if reverse else 0, fv.shape[0])
    return out


def galerkin(shape, A):
    """P^T A P of an ELL-27 operator onto the lattice of half the edges."""
    nz, ny, nx = shape
    Ac = np.zeros(27 * (nx // 2) * (ny // 2) * (nz // 2))
    lib().orc_galerkin_ell(nx, ny, nz, _d(A).reshape(-1), Ac)
    return Ac.reshape(27, -1)


def mg_homogenize(kappa, rtol=1e-10, want_chi=False, want_S=False, fp32=False):
    """kappa^H by MG-preconditioned CG (FP64; fp32=True: the FP32 CPU port).
    Returns (C, chi or None, S or None, its[3], relres[3])."""
    k = _d(kappa)
    nz, ny, nx = k.shape
    n = k.size
    C = np.zeros(9)
    chi = np.zeros(3 * n) if want_chi else None
    S = np.zeros(6 * n) if want_S else None
    its = np.zeros(3, np.int32)
    rr = np.zeros(3)
    f = lib().orc32_mg_homogenize if fp32 else lib().orc_mg_homogenize
    st = f(nx, ny, nz, k.reshape(-1), rtol, chi.ctypes.data_as(ctypes.c_void_p) if want_chi else None, C,
           S.ctypes.data_as(ctypes.c_void_p) if want_S else None, its, rr)
    if st != 0:
        raise RuntimeError(f"orc_mg_homogenize failed ({st})")
    return (C.reshape(3, 3), None if chi is None else chi.reshape(3, n), None if S is None else S.reshape(6, n),
            its.tolist(), rr.tolist())


def element_energy(shape, chi):
    nz, ny, nx = shape
    S = np.zeros(6 * nx * ny * nz)
    lib().orc_element_energy(nx, ny, nz, _d(chi).reshape(-1), S)
    return S.reshape(6, -1)


def sensitivity(S6, dC):
    S = _d(S6)
    n = S.shape[1]
    out = np.zeros(n)
    lib().orc_sensitivity(n, S.reshape(-1), _d(dC).reshape(-1), out)
    return out


def filter_fwd(x, kind=0, r=1.5, periodic=False, want_W=False):
    a = _d(x)
    nz, ny, nx = a.shape
    out = np.zeros(a.size)
    W = np.zeros(a.size) if want_W else None
    lib().orc_filter_fwd(nx, ny, nz, kind, r, 1 if periodic else 0, a.reshape(-1), out,
                         W.ctypes.data_as(ctypes.c_void_p) if want_W else None)
    out = out.reshape(a.shape)
    return (out, W.reshape(a.shape)) if want_W else out


def filter_adj(x, kind=0, r=1.5, periodic=False):
    a = _d(x)
    nz, ny, nx = a.shape
    out = np.zeros(a.size)
    lib().orc_filter_adj(nx, ny, nz, kind, r, 1 if periodic else 0, a.reshape(-1), out)
    return out.reshape(a.shape)


def oc_filter_sens(rho, sens, r=2.0):
    a = _d(rho)
    nz, ny, nx = a.shape
    out = np.zeros(a.size)
    lib().orc_oc_filter_sens(nx, ny, nz, r, a.reshape(-1), _d(sens).reshape(-1), out)
    return out.reshape(a.shape)


def oc_update(sens, rho, volratio, min_rho=0.001, step=0.02, damp=0.5, max_steps=20, vol_tol=1e-4):
    s = _d(sens).reshape(-1)
    r = _d(rho).reshape(-1)
    out = np.zeros(r.size)
    g = ctypes.c_double()
    steps = lib().orc_oc_update(r.size, s, r, out, volratio, min_rho, step, damp, max_steps, vol_tol,
                                ctypes.byref(g))
    return out, steps, g.value


class MMA:
    """State of the oracle's MMA (xold1, xold2, low, upp), like MMAContext."""

    def __init__(self, m, n, a0=1.0, a=0.0, c=1e3, d=1.0, xmin=0.0, xmax=1.0):
        self.m, self.n = m, n
        self.a0 = a0
        self.a = np.full(m, float(a))
        self.c = np.full(m, float(c))
        self.d = np.full(m, float(d))
        self.xmin = np.broadcast_to(np.asarray(xmin, np.float64), (n,)).copy()
        self.xmax = np.broadcast_to(np.asarray(xmax, np.float64), (n,)).copy()
        self.xold1 = np.zeros(n)
        self.xold2 = np.zeros(n)
        self.low = np.zeros(n)
        self.upp = np.zeros(n)
        self.move = 1.0
        self.epsimin = 1e-7  # the reference's last barrier level (mmaOpt.cu:846)

    @staticmethod
    def stats(reset=False):
        """(Newton steps, steps whose length the design variables' bounds set)
        over the subsolver runs since the last reset."""
        a, b = ctypes.c_long(), ctypes.c_long()
        lib().orc_mma_stats(ctypes.byref(a), ctypes.byref(b))
        if reset:
            lib().orc_mma_stats_reset()
        return a.value, b.value

    def update(self, itn, x, df0dx, gval, dgdx):
        """Returns the new x (float64) and the Newton step count."""
        xv = _d(x).reshape(-1).copy()
        its = lib().orc_mmasub(self.m, self.n, itn, xv, self.xmin, self.xmax, self.xold1, self.xold2,
                               _d(df0dx).reshape(-1), _d(gval).reshape(-1), _d(dgdx).reshape(-1),
                               self.low, self.upp, self.a0, self.a, self.c, self.d, self.move, self.epsimin)
        return xv, its


# ---- symmetrisation ---------------------------------------------------------
# symmetrizeField (Framework.cuh:221-235) with symetrize_tensor_kernel
# (TensorExpression.cuh:70-300): every value becomes the mean over its orbit
# under a group of the cube's symmetries about the cell centre. A group
# element (perm, flip) maps coordinate c to c' with c'[j] = c[perm[j]],
# reflected to n - 1 - c'[j] when bit j of flip is set.
_PERMS = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]  # 3 even, 3 odd


def sym_group(sym):
    """Group elements of 'reflect3' (8 reflections), 'reflect6' (48:
    reflections x permutations), 'rotate3' (the 24 rotations: permutation
    parity equals reflection-count parity) and 'rotate2' (point reflection)."""
    if sym == "reflect3":
        return [(_PERMS[0], f) for f in range(8)]
    if sym == "reflect6":
        return [(p, f) for p in _PERMS for f in range(8)]
    if sym == "rotate3":
        return [(p, f) for i, p in enumerate(_PERMS) for f in range(8) if bin(f).count("1") % 2 == (1 if i >= 3 else 0)]
    if sym == "rotate2":
        return [(_PERMS[0], 0), (_PERMS[0], 7)]
    raise ValueError(sym)


def sym_apply(x, elem):
    """x o g for one group element: y[c] = x[g(c)], arrays (nz, ny, nx)."""
    perm, flip = elem
    nz, ny, nx = x.shape
    dims = (nx, ny, nz)
    c = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")  # c[j][x, y, z]
    q = []
    for j in range(3):
        v = c[perm[j]]
        q.append(dims[j] - 1 - v if (flip >> j) & 1 else v)
    return np.transpose(x[q[2], q[1], q[0]], (2, 1, 0))


def symmetrize(x, sym):
    """Orbit mean of x (FP64)."""
    x = np.asarray(x, np.float64)
    G = sym_group(sym)
    return sum(sym_apply(x, g) for g in G) / len(G)
