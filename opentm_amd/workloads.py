"""Seeded synthetic densities for the benchmark configurations.

All generators use numpy's counter-based Philox bit generator, so a seed
fixes the field on any machine. Fields are float32 arrays of shape
(nz, ny, nx) (x fastest in memory).
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def _wavevectors(rng, count, kmax=3):
    """`count` non-zero integer wavevectors with Euclidean norm <= kmax."""
    pool = [(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1)
            for c in range(-kmax, kmax + 1) if 0 < a * a + b * b + c * c <= kmax * kmax]
    idx = rng.choice(len(pool), size=count, replace=False)
    return np.array([pool[i] for i in idx], dtype=np.int64)


def cos_modes(n, seed, modes=8, kmax=3):
    """sum_m a_m cos(2 pi k_m . x / n + phi_m) on an n^3 grid (x = voxel index),
    with a_m ~ N(0,1)/8 and phi_m ~ U[0, 2 pi)."""
    rng = _rng(seed)
    k = _wavevectors(rng, modes, kmax)
    a = rng.standard_normal(modes) / 8.0
    phi = rng.uniform(0.0, 2.0 * np.pi, modes)
    x = np.arange(n, dtype=np.float64)
    field = np.zeros((n, n, n), dtype=np.float64)
    for m in range(modes):
        # separable product of complex exponentials keeps memory at O(n^3)
        ex = np.exp(2j * np.pi * k[m, 0] * x / n)
        ey = np.exp(2j * np.pi * k[m, 1] * x / n)
        ez = np.exp(2j * np.pi * k[m, 2] * x / n)
        field += a[m] * np.real(np.exp(1j * phi[m]) * ez[:, None, None] * ey[None, :, None] * ex[None, None, :])
    return field


def plain_density(n, seed, mean=0.3):
    """Density of the plain-method configuration (32^3 and 256^3 cells):
    clamp(mean + 0.25 * cos-mode field, [1e-3, 1])."""
    rho = mean + 0.25 * cos_modes(n, seed)
    return np.clip(rho, 1e-3, 1.0).astype(np.float32)


def oc_start_density(n, seed, mean=0.3, amplitude=0.05):
    """Start of the OC loop: mean + amplitude * (cos-mode field / max |field|)."""
    f = cos_modes(n, seed)
    f = f / max(np.abs(f).max(), 1e-30)
    return np.clip(mean + amplitude * f, 1e-3, 1.0).astype(np.float32)


def simp(rho, k0=1.0, k1=1e-3, penal=3):
    """kappa = k1 + (k0 - k1) rho^p (BASELINE material law)."""
    r = np.asarray(rho, dtype=np.float64)
    return (k1 + (k0 - k1) * r ** penal).astype(np.float32)


def grf_density(n, seed, corr_len, vol, beta=8.0):
    """Serving-form density: a periodic Gaussian random field with Gaussian
    correlation of length `corr_len` voxels (unit variance), passed through
    the smooth projection 1 / (1 + exp(-beta (phi - t))) with the threshold
    t bisected so the mean density is `vol`; clamped to [1e-3, 1]."""
    rng = _rng(seed)
    white = rng.standard_normal((n, n, n))
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    kr = 2.0 * np.pi * np.fft.rfftfreq(n)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kr[None, None, :] ** 2
    phi = np.fft.irfftn(np.fft.rfftn(white) * np.exp(-0.5 * k2 * corr_len ** 2), s=(n, n, n), axes=(0, 1, 2))
    phi = (phi / max(phi.std(), 1e-30)).astype(np.float32)
    lo, hi = -10.0, 10.0
    for _ in range(32):
        t = 0.5 * (lo + hi)
        m = np.mean(1.0 / (1.0 + np.exp(np.float32(-beta) * (phi - np.float32(t)))), dtype=np.float64)
        if m > vol:
            lo = t
        else:
            hi = t
    rho = 1.0 / (1.0 + np.exp(np.float32(-beta) * (phi - np.float32(0.5 * (lo + hi)))))
    return np.clip(rho, 1e-3, 1.0).astype(np.float32)
