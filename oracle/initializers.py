"""The reference's random start designs restated in numpy (FP64), the
checker for the engine's otm_init_design.

TEST INFRASTRUCTURE ONLY.

- rand_tri: randTri (Framework.cuh:238-310) with the basis of
  randTribase_sincos_kernel (Framework.cuh:106-214), the level-set
  bisection of sigmoid(., 15, t) (lib.cuh:1381) and the final map.
- noise: Framework.cu:75-88 with projectTensor_kernel
  (TensorExpression.cu:106-115), on the engine's counter-based uniforms.
- init_density_tail: symmetrise and clamp to [0.1, 1] (Framework.cu:140-144).
"""
import numpy as np

from .oracle import symmetrize

PI2 = np.float64(np.float32(3.1415926) * np.float32(2.0))


def basis(shape, P, centered=False, quat=None):
    """psi[a*P + f] = cos(2 pi (f+1) p_a), psi[3P + a*P + f] = sin(...), with
    p = index / (edge - 1) per axis (x fastest); shapes (6P, nz, ny, nx)."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [x / (nx - 1.0), y / (ny - 1.0), z / (nz - 1.0)]
    if centered:
        w, qx, qy, qz = np.asarray(quat, np.float64) / np.linalg.norm(quat)
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - w * qz), 2 * (qx * qz + w * qy)],
                      [2 * (qx * qy + w * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - w * qx)],
                      [2 * (qx * qz - w * qy), 2 * (qy * qz + w * qx), 1 - 2 * (qx * qx + qy * qy)]])
        v = [c - 0.5 for c in p]
        p = [R[i, 0] * v[0] + R[i, 1] * v[1] + R[i, 2] * v[2] for i in range(3)]
    psi = np.zeros((6 * P,) + tuple(shape))
    for a in range(3):
        for f in range(P):
            psi[a * P + f] = np.cos(PI2 * (f + 1) * p[a])
            psi[3 * P + a * P + f] = np.sin(PI2 * (f + 1) * p[a])
    return psi


def rand_tri_field(shape, P, coeffs, centered=False):
    """The basis sum of randTribase_sincos_kernel: first order
    c_i C[i%3][i/3] (i < 3P) and c_{3P+i} S[i%3][i/3]; second order
    c_{6P + 6P j + k} psi_j psi_k."""
    c = np.asarray(coeffs, np.float64)
    nb1 = 6 * P
    psi = basis(shape, P, centered, c[nb1 + 36 * P * P:] if centered else None)
    s = np.zeros(shape)
    for i in range(3 * P):
        s += c[i] * psi[(i % 3) * P + i // 3]
        s += c[3 * P + i] * psi[3 * P + (i % 3) * P + i // 3]
    C2 = c[nb1:nb1 + nb1 * nb1].reshape(nb1, nb1)
    s += np.einsum("j...,jk,k...->...", psi, C2, psi)
    return s


def level_map(val, t, hi):
    return 1.0 / (1.0 + np.exp(-15.0 * (val - t))) * (hi - 0.001) + 0.001


def rand_tri(shape, P, coeffs, volratio, centered=False, sym="NONE"):
    """randTri followed by initDensity's symmetrise + clamp(0.1, 1)."""
    v = rand_tri_field(shape, P, coeffs, centered)
    if sym != "NONE":
        v = symmetrize(v, sym)
    V = np.float32(volratio)
    hi = min(float(V) * 1.5, 1.0)
    tu = np.float32(P * P * 37.0)
    tl = -tu
    t = (tu + tl) / np.float32(2)
    for _ in range(60):
        t = (tu + tl) / np.float32(2)
        s = np.float32(level_map(v, float(t), hi).mean())
        if abs(s - V) / V < np.float32(1e-3):
            break
        if s > V:
            tl = t
        if s < V:
            tu = t
    return init_density_tail(level_map(v, float(t), hi), sym), float(t)


def hash_uniform(seed, n):
    """The engine's counter-based uniforms: top 24 bits of splitmix64 of
    (seed << 32) + index."""
    z = (np.uint64(seed) << np.uint64(32)) + np.arange(n, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) / 16777216.0


def tanproj(v, beta=20.0, tau=0.5):
    """lib.cuh:1372-1379."""
    t = np.tanh(beta * tau)
    return (t + np.tanh(beta * (v - tau))) / (t + np.tanh(beta * (1.0 - tau)))


def noise(shape, seed, volratio, sym="NONE"):
    n = int(np.prod(shape))
    v = hash_uniform(seed, n).reshape(shape)
    if sym != "NONE":
        v = symmetrize(v, sym)
    r = tanproj(v)
    r = np.where(r < 0.5, 1e-9, np.where(r > 0.5, 1.0, r))
    s = volratio / r.mean()
    r = np.clip(r * s, 0.001, 1.0)
    return init_density_tail(r, sym)


def init_density_tail(v, sym):
    if sym != "NONE":
        v = symmetrize(v, sym)
    return np.clip(v, 0.1, 1.0)
