"""Seeded synthetic workloads (trajectories and images), pure numpy.

The same generator runs in the bench, the tests and the fixture writer, so a
fixture written here and a test run on another machine see the same inputs.
Sizes and seeds:

* ``plain2d``  : Nd 64^2, Kd 128^2, Jd 6^2, M 4096, om ~ U[-pi,pi)^2 (seed 101),
                 x = N(0,1) + i N(0,1) (seed 102).
* ``radial2d`` : Nd 256^2, Kd 512^2, Jd 6^2, golden-angle radial trajectory,
                 402 spokes x 512 readout samples (M 205 824); image = sum of
                 10 random ellipses (seed 201); 1 % complex noise (seed 202).
* ``uniform3d``: Nd 128^3, Kd 256^3, Jd 6^3, M 2 097 152, om ~ U[-pi,pi)^3
                 (seed 301), x = N(0,1) + i N(0,1) (seed 302).
"""
import numpy as np

GOLDEN_RATIO_ANGLE = np.pi * (np.sqrt(5.0) - 1.0) / 2.0


def uniform_om(M, ndims, seed):
    """M coordinates i.i.d. uniform in [-pi, pi)^ndims, float64."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-np.pi, np.pi, size=(M, ndims))


def gaussian_image(Nd, seed, batch=None):
    """complex64 image, real and imaginary parts i.i.d. N(0, 1)."""
    rng = np.random.default_rng(seed)
    shape = tuple(Nd) + ((batch,) if batch else ())
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return x.astype(np.complex64)


def radial_om(spokes, readout):
    """Golden-angle radial trajectory (spokes*readout, 2), radius in [-pi, pi).

    Spoke s has angle s * pi (sqrt5 - 1) / 2; sample i has radius
    pi (i - readout/2) / (readout/2).
    """
    theta = np.arange(spokes) * GOLDEN_RATIO_ANGLE
    half = readout // 2
    r = np.pi * (np.arange(readout) - half) / half
    kx = np.outer(np.cos(theta), r).ravel()
    ky = np.outer(np.sin(theta), r).ravel()
    return np.stack([kx, ky], axis=1)


def spiral_om(M, N, seed, turns=None):
    """Variable-density spiral, sample density ~ 1/(1 + r^2); (M, 2)."""
    rng = np.random.default_rng(seed)
    turns = turns or max(8, N // 8)
    t = np.linspace(0.0, 1.0, M, endpoint=False)
    # radius r(t) with dr/dt ~ (1 + r^2): r = tan(t * atan(pi))
    r = np.tan(t * np.arctan(np.pi))
    phase = rng.uniform(0, 2 * np.pi)
    ang = 2 * np.pi * turns * t + phase
    om = np.stack([r * np.cos(ang), r * np.sin(ang)], axis=1)
    return np.clip(om, -np.pi, np.pi - 1e-9)


def ellipse_phantom(Nd, seed, count=10):
    """Sum of ``count`` random ellipses on a grid over [-1, 1)^2, complex64.

    centres U(-0.6, 0.6), semi-axes U(0.05, 0.4), angles U(0, pi),
    intensities U(-1, 1).  Axis 0 is y, axis 1 is x.
    """
    rng = np.random.default_rng(seed)
    yy = -1.0 + 2.0 * np.arange(Nd[0]) / Nd[0]
    xx = -1.0 + 2.0 * np.arange(Nd[1]) / Nd[1]
    Y, X = np.meshgrid(yy, xx, indexing='ij')
    img = np.zeros(Nd, dtype=np.float64)
    for _ in range(count):
        cx, cy = rng.uniform(-0.6, 0.6, size=2)
        a, b = rng.uniform(0.05, 0.4, size=2)
        phi = rng.uniform(0, np.pi)
        val = rng.uniform(-1, 1)
        c, s = np.cos(phi), np.sin(phi)
        xr = (X - cx) * c + (Y - cy) * s
        yr = -(X - cx) * s + (Y - cy) * c
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += val
    return img.astype(np.complex64)


def complex_noise(y, level, seed):
    """Complex Gaussian noise with rms = level * rms(y)."""
    rng = np.random.default_rng(seed)
    rms = np.sqrt(np.mean(np.abs(y) ** 2))
    n = rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape)
    return (level * rms / np.sqrt(2.0)) * n


def plain2d():
    Nd, Kd, Jd = (64, 64), (128, 128), (6, 6)
    return dict(om=uniform_om(4096, 2, 101), x=gaussian_image(Nd, 102),
                Nd=Nd, Kd=Kd, Jd=Jd)


def radial2d():
    Nd, Kd, Jd = (256, 256), (512, 512), (6, 6)
    return dict(om=radial_om(402, 512), x=ellipse_phantom(Nd, 201),
                Nd=Nd, Kd=Kd, Jd=Jd, noise_seed=202)


def uniform3d(M=2097152):
    Nd, Kd, Jd = (128, 128, 128), (256, 256, 256), (6, 6, 6)
    return dict(om=uniform_om(M, 3, 301), x=gaussian_image(Nd, 302),
                Nd=Nd, Kd=Kd, Jd=Jd)


def fingerprint(a):
    """Order-sensitive float64 checksum of an array (for fixture checks)."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        a = np.stack([a.real, a.imag], axis=-1)
    a = a.astype(np.float64).ravel()
    w = np.cos(np.arange(a.size) * 0.7071067811865476)
    return float(np.dot(a, w)), float(np.dot(a, a))


def batched2d(n=64, seed0=400):
    """``n`` independent 2-D problems of uneven size (the batched workload).

    Problem i (seed 400 + i): Nd = (N, N) with N drawn from {128, 192, 256},
    Kd = 2 Nd, Jd = (6, 6); trajectory either golden-angle radial with
    64-402 spokes x 2N readout samples, or a variable-density spiral with
    density ~ 1/(1 + r^2) and M drawn in [16384, 205824]; image = 10-ellipse
    phantom.
    """
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        N = int(rng.choice([128, 192, 256]))
        if rng.uniform() < 0.5:
            spokes = int(rng.integers(64, 403))
            om = radial_om(spokes, 2 * N)
            kind = 'radial'
        else:
            M = int(rng.integers(16384, 205825))
            om = spiral_om(M, N, seed0 + i)
            kind = 'spiral'
        out.append(dict(om=om, Nd=(N, N), Kd=(2 * N, 2 * N), Jd=(6, 6),
                        x=ellipse_phantom((N, N), seed0 + i), kind=kind))
    return out


def coil_maps(Nd, ncoil, seed):
    """Smooth complex coil sensitivities: degree-2 polynomials in (y, x) over
    [-1, 1)^2 with random complex coefficients, normalised so that
    sum_c |S_c|^2 = 1 at every pixel.  Shape Nd + (ncoil,), complex64."""
    rng = np.random.default_rng(seed)
    yy = -1.0 + 2.0 * np.arange(Nd[0]) / Nd[0]
    xx = -1.0 + 2.0 * np.arange(Nd[1]) / Nd[1]
    Y, X = np.meshgrid(yy, xx, indexing='ij')
    basis = np.stack([np.ones_like(X), X, Y, X * X, X * Y, Y * Y], axis=-1)
    coef = rng.standard_normal((6, ncoil)) + 1j * rng.standard_normal((6, ncoil))
    coef[0] += 3.0  # keep every map away from zero
    S = basis @ coef
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=-1, keepdims=True))
    return S.astype(np.complex64)


def multicoil2d(perturb=1e-3, seed=501):
    """Differentiable multi-coil workload: Nd 256^2, Kd 512^2, Jd 6^2, 256
    radial spokes x 512 samples (M 131 072), 8 coils, ellipse image; the
    data trajectory is om + N(0, perturb) rad."""
    rng = np.random.default_rng(seed)
    om = radial_om(256, 512)
    om_data = np.clip(om + rng.normal(0.0, perturb, om.shape), -np.pi, np.pi - 1e-9)
    Nd = (256, 256)
    return dict(om=om, om_data=om_data, Nd=Nd, Kd=(512, 512), Jd=(6, 6),
                x=ellipse_phantom(Nd, seed + 1), S=coil_maps(Nd, 8, seed + 2))
