"""GPU parity of the preconditioner's parts, compared directly (PCG reaches
the right kappa^H with any SPD preconditioner, so end-to-end parity cannot
catch a mistake in them): transfers against the dense trilinear P, exact and
brick Gauss-Seidel sweeps against the oracle's CPU sweeps in the reference's
colour order, the coarsest solve against the pseudo-inverse of the coarse
operator. FP32 kernels against FP64 references, held to 2e-6 relative.
"""
import sys

import numpy as np
import pytest

import opentm_amd as otm
from opentm_amd import workloads as wl

from test_oracle_mg import dense, prolongation

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(reso, k1=1e-3, **kw):
    cfg = otm.HomoConfig(reso, (1.0, k1))
    for key, v in kw.items():
        setattr(cfg, key, v)
    return otm.Homogenization_H(cfg)


def shape_of(level_reso):
    nx, ny, nz = level_reso
    return (nz, ny, nx)


@pytest.mark.parametrize("reso", [(16, 16, 16), (16, 8, 12), (128, 8, 8)])
def test_transfers_against_dense_P(reso):
    """Restriction and prolongation against the dense trilinear P; (128, 8, 8)
    has 64-node coarse rows, where the restriction reads fine pairs and takes
    the left neighbour from the lane before (its lane 0 wraps to x = nx-1
    for the first wave of a row)."""
    say("transfers", reso)
    rng = np.random.default_rng(21)
    nx, ny, nz = reso
    hom = make(reso)
    hom.update(rng.uniform(1e-3, 1.0, (nz, ny, nx)).astype(np.float32))
    lv = hom.levels()
    P = prolongation(shape_of(lv[0]))
    nf, nc = P.shape
    r = rng.standard_normal((3, nf)).astype(np.float32)
    fc = hom.restrict(1, r).reshape(3, nc)
    for k in range(3):
        assert rel(fc[k], P.T @ r[k].astype(np.float64)) < 2e-6
    uc = rng.standard_normal((3, nc)).astype(np.float32)
    uf = rng.standard_normal((3, nf)).astype(np.float32)
    out = hom.prolong_add(1, uc, uf).reshape(3, nf)
    for k in range(3):
        assert rel(out[k], uf[k] + P @ uc[k].astype(np.float64)) < 2e-6
    # P 1 = 1: the constant is prolongated exactly
    one = hom.prolong_add(1, np.ones(3 * nc, np.float32), np.zeros(3 * nf, np.float32))
    assert np.all(one == 1.0)


def level_operator(orc, hom, level, kappa):
    """ELL-27 operator of a level: the oracle's assembly on the fine level,
    the GPU's own stencil on a coarse one (pinned against the dense P^T K P
    by test_gpu_parity::test_galerkin_coarse_operator)."""
    if level == 0:
        return orc.assemble(kappa.astype(np.float64) + 1e-9)
    return hom.stencil(level).astype(np.float64)


@pytest.mark.parametrize("reso,level", [((16, 16, 16), 0), ((16, 8, 12), 0), ((32, 32, 32), 1), ((24, 24, 24), 1),
                                        ((32, 16, 24), 1), ((32, 32, 32), 2)])
def test_exact_colour_sweeps_match_oracle(orc, reso, level):
    """Exact multicolour Gauss-Seidel: the engine's reference-order sweep
    (mode 1) equals the oracle's sweep in the reference's colour order
    0..7, its forward sweep (mode 0) the oracle's reverse; two sweeps from
    a random start, three right-hand sides."""
    say("exact sweeps", reso, "level", level)
    rng = np.random.default_rng(22)
    nx, ny, nz = reso
    k = rng.uniform(1e-3, 1.0, (nz, ny, nx)).astype(np.float32)
    hom = make(reso)
    hom.update(k)
    shape = shape_of(hom.levels()[level])
    n = int(np.prod(shape))
    A = level_operator(orc, hom, level, k)
    f = rng.standard_normal((3, n)).astype(np.float32)
    u = rng.standard_normal((3, n)).astype(np.float32)
    for mode, reverse in ((1, False), (0, True)):
        g = hom.smooth_level(level, f, u, sweeps=2, mode=mode).reshape(3, n)
        o = orc.gs(shape, A, f, u, sweeps=2, reverse=reverse)
        errs = [rel(g[c], o[c]) for c in range(3)]
        say("  mode", mode, "max rel err", max(errs))
        assert max(errs) < 2e-6


@pytest.mark.parametrize("reso,level,brick", [
    ((64, 8, 16), 0, (32, 4, 8)),     # fine level: 2 x 2 x 2 tiles of 32 x 4 x 8
    ((48, 12, 8), 0, (16, 4, 8)),     # 16-wide fine tiles
    ((32, 32, 32), 1, (8, 8, 8)),     # coarse 16^3: 2 x 2 x 2 bricks
    ((24, 24, 24), 1, (8, 8, 8)),     # coarse 12^3: bricks cut at the edge
    ((16, 16, 16), 1, (8, 8, 8))])    # coarse 8^3: one brick, periodic seam frozen
def test_brick_smoother_matches_block_restatement(orc, reso, level, brick):
    """The preconditioner's brick smoother against the oracle's block
    Gauss-Seidel (exact inside each brick, couplings leaving it frozen at the
    entry values), both colour orders, two sweeps."""
    say("brick sweeps", reso, "level", level, "brick", brick)
    rng = np.random.default_rng(23)
    nx, ny, nz = reso
    k = rng.uniform(1e-3, 1.0, (nz, ny, nx)).astype(np.float32)
    hom = make(reso)
    hom.update(k)
    shape = shape_of(hom.levels()[level])
    n = int(np.prod(shape))
    A = level_operator(orc, hom, level, k)
    f = rng.standard_normal((3, n)).astype(np.float32)
    u = rng.standard_normal((3, n)).astype(np.float32)
    for mode, reverse in ((3, False), (2, True)):
        g = hom.smooth_level(level, f, u, sweeps=2, mode=mode).reshape(3, n)
        o = orc.brick_gs(shape, A, f, u, brick, sweeps=2, reverse=reverse)
        errs = [rel(g[c], o[c]) for c in range(3)]
        say("  mode", mode, "max rel err", max(errs))
        assert max(errs) < 2e-6
        # and it is not the exact sweep: the frozen couplings matter
        if level == 0 or shape[0] > 8:
            ex = orc.gs(shape, A, f, u, sweeps=2, reverse=reverse)
            assert rel(g[0], ex[0]) > 1e-4


@pytest.mark.parametrize("reso", [16, 48])
def test_coarsest_solve_is_the_pseudo_inverse(reso):
    """u = B^{-1} f on the coarsest level (n = 64: inverse formed in LDS;
    n = 216: the blocked global-memory Gauss-Jordan) equals the minimum-norm
    solution K^+ f of the singular periodic coarse operator for mean-free f
    (numpy's dense lstsq), the regularisation fixing only the constant."""
    say("coarsest", reso)
    rng = np.random.default_rng(24)
    hom = make(reso)
    hom.update(wl.simp(wl.plain_density(reso, 2)))
    L = len(hom.levels()) - 1
    shape = shape_of(hom.levels()[L])
    n = int(np.prod(shape))
    K = dense(hom.stencil(L).astype(np.float64), shape)
    K = 0.5 * (K + K.T)
    # the FP32 stencil's rows sum to zero only to rounding (about 4e-8 of the
    # largest entry): lstsq on that K sees no null space and its answer
    # carries a large constant, so the reference operator is the exactly
    # singular one: off-diagonals as stored, diagonal = -(their row sum)
    rowsum = float(np.abs(K.sum(axis=1)).max() / np.abs(K).max())
    say("  row sums", rowsum)
    assert rowsum < 1e-6
    K0 = K - np.diag(np.diag(K))
    K0 -= np.diag(K0.sum(axis=1))
    f = rng.standard_normal((3, n))
    f -= f.mean(axis=1, keepdims=True)
    u = hom.coarsest_solve(f.astype(np.float32)).reshape(3, n).astype(np.float64)
    for c in range(3):
        ref = np.linalg.lstsq(K0, f[c], rcond=None)[0]
        ref -= ref.mean()
        err = rel(u[c] - u[c].mean(), ref)
        res = rel(K0 @ u[c], f[c])
        say("  rhs", c, "n", n, "err", err, "residual", res)
        assert err < 1e-5
        assert res < 1e-5


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_fine_tile_variants_match_block_restatement(orc, variant):
    """Every fine smoother tile shape the autotuner can pick is the oracle's
    block Gauss-Seidel with that tile, both colour orders, two sweeps."""
    shape3 = otm.fine_tile_shape(variant)[:3]
    say("fine tile variant", variant, shape3)
    rng = np.random.default_rng(25)
    reso = (64, 16, 16)
    k = rng.uniform(1e-3, 1.0, (16, 16, 64)).astype(np.float32)
    hom = make(reso, fine_tile=variant)
    assert hom.fine_tile()[0] == variant
    hom.update(k)
    shape = shape_of(hom.levels()[0])
    n = int(np.prod(shape))
    A = orc.assemble(k.astype(np.float64) + 1e-9)
    f = rng.standard_normal((3, n)).astype(np.float32)
    u = rng.standard_normal((3, n)).astype(np.float32)
    for mode, reverse in ((3, False), (2, True)):
        g = hom.smooth_level(0, f, u, sweeps=2, mode=mode).reshape(3, n)
        o = orc.brick_gs(shape, A, f, u, shape3, sweeps=2, reverse=reverse)
        errs = [rel(g[c], o[c]) for c in range(3)]
        say("  mode", mode, "max rel err", max(errs))
        assert max(errs) < 2e-6


def test_autotune_picks_a_candidate_that_matches_the_oracle(orc):
    """The autotuner times every admissible (tile shape, coarse sweeps)
    candidate; the context keeps the fastest, its solve matches the FP64
    oracle, and a context pinned to the same choice solves bitwise alike."""
    n = 32
    k = wl.simp(wl.plain_density(n, 5))
    hom = make(n)
    hom.update(k)
    rows, chosen = hom.autotune()
    say("autotune", [(r["fine_tile"], r["coarse_sweeps"], round(r["ms"], 3), r["iters"], r["ok"]) for r in rows],
        "chosen", chosen)
    assert len(rows) == 10 and 0 <= chosen < len(rows) and rows[chosen]["ok"]
    assert rows[chosen]["ms"] == min(r["ms"] for r in rows if r["ok"])
    assert hom.fine_tile()[0] == rows[chosen]["fine_tile"]
    C = hom.heatMatrix()
    Co = orc.homogenize(k.astype(np.float64), rtol=1e-12)[0]
    assert rel(C, Co) < 1e-5
    pinned = make(n, fine_tile=rows[chosen]["fine_tile"], coarse_sweeps=rows[chosen]["coarse_sweeps"])
    pinned.update(k)
    assert np.array_equal(pinned.heatMatrix(), C)
    # warm: every candidate starts from the solution just computed and the
    # iterate the context keeps afterwards is that solution
    rows_w, chosen_w = hom.autotune(warm=True)
    assert all(r["ok"] for r in rows_w) and 0 <= chosen_w < len(rows_w)
    assert rel(hom.heatMatrix(), Co) < 1e-5
    for h in (hom, pinned):
        h.close()


@pytest.mark.parametrize("refine", [0, 1])
@pytest.mark.parametrize("reso", [(24, 8, 12), (40, 16, 8), (256, 8, 8), (384, 8, 8), (1000, 16, 16)])
def test_start_residual_norms_match_oracle(orc, reso, refine):
    """The FP64 start residual r = b - K x that every solve and restart
    begins from (by whole x-rows for rows of up to 256 nodes: 24 to 256
    here, one to ten row groups per workgroup, segments that cross z-planes;
    pointwise for the 384- and 1000-node rows): the relative norms
    ||r_k|| / ||b_k|| that otm_mg_cycles reports before its cycle equal the
    oracle's assembly applied to the same iterate, from the FP32 iterate
    (plain form) and from its FP64 copy (refinement form). The iterate is
    scaled so that K x and b are of the same size and both terms count."""
    say("start residual", reso, "refine", refine)
    rng = np.random.default_rng(26)
    nx, ny, nz = reso
    shape = (nz, ny, nx)
    n = nx * ny * nz
    k = rng.uniform(1e-3, 1.0, shape).astype(np.float32)
    hom = make(reso, fp64_refine=refine)
    hom.update(k)
    keff = (k + np.float32(1e-9)).astype(np.float64)  # the context's floor, added in FP32
    A = orc.assemble(keff)
    u = rng.standard_normal((3, n))
    b = [orc.rhs(keff, c) for c in range(3)]
    for c in range(3):
        u[c] *= np.linalg.norm(b[c]) / np.linalg.norm(orc.matvec(shape, A, u[c]))
    u = u.astype(np.float32)
    hom.set_chi(u)
    got = hom.mg_cycles(1)
    for c in range(3):
        r = b[c] - orc.matvec(shape, A, u[c].astype(np.float64))
        want = np.linalg.norm(r) / np.linalg.norm(b[c])
        say("  case", c, "relres", got[c], "oracle", want)
        assert 0.5 < want < 3.0
        assert abs(got[c] - want) < 1e-6 * want
    hom.close()
