"""Seeded synthetic SPECT workloads (ellipsoid phantoms, attenuation maps,
depth-dependent Gaussian PSFs, camera sets).

Arrays follow the numpy conventions of niftyrec_amd.api: volumes (nz, ny, nx)
with y the axial (rotation) axis, psf (nz, K, K), cameras (V,) radians about y.
The reference has no phantom generator for this path (its et_array_make_psf
is a stub, _et_array_interface.cpp:896-920), so the recipe is this project's.
"""
import numpy as np


def ellipsoid_mask(shape, center, semi_axes):
    """Boolean mask of an axis-aligned ellipsoid; center/semi_axes in (x, y, z)."""
    nz, ny, nx = shape
    z, y, x = np.ogrid[:nz, :ny, :nx]
    cx, cy, cz = center
    ax, ay, az = semi_axes
    return ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 + ((z - cz) / az) ** 2 <= 1.0


def gaussian_psf(n_planes, K, sigma_near, sigma_far):
    """(n_planes, K, K) unit-sum Gaussians, sigma linear in depth."""
    r = (K - 1) / 2.0
    t = np.arange(K) - r
    out = np.empty((n_planes, K, K), dtype=np.float64)
    for d in range(n_planes):
        s = sigma_near + (sigma_far - sigma_near) * (d / max(n_planes - 1, 1))
        g = np.exp(-0.5 * (t / s) ** 2)
        out[d] = np.outer(g, g)
        out[d] /= out[d].sum()
    return out.astype(np.float32)


def phantom(shape, n_lesions, seed, body_axes=None, lesion_axes=(2.0, 8.0), lesion_values=(2.0, 10.0),
            mu_body=0.15, extra_mu=()):
    """Activity and attenuation of an ellipsoid body with random lesions.

    extra_mu: sequence of (center_fraction(x,y,z), semi_axes, value) ellipsoids
    whose attenuation replaces the body value inside them."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    c = ((nx - 1) / 2.0, (ny - 1) / 2.0, (nz - 1) / 2.0)
    if body_axes is None:
        body_axes = (0.4375 * nx, 0.375 * ny, 0.4375 * nz)
    body = ellipsoid_mask(shape, c, body_axes)
    act = body.astype(np.float64)
    for _ in range(n_lesions):
        ax = rng.uniform(*lesion_axes, size=3)
        # centre inside the body (rejection sampling on the body ellipsoid)
        while True:
            p = rng.uniform(-1, 1, size=3)
            if (p ** 2).sum() <= 0.6:
                break
        ctr = tuple(c[i] + p[i] * body_axes[i] for i in range(3))
        act += rng.uniform(*lesion_values) * ellipsoid_mask(shape, ctr, ax)
    mu = np.where(body, mu_body, 0.0)
    for frac, axes, val in extra_mu:
        ctr = tuple(c[i] + frac[i] * body_axes[i] for i in range(3))
        mu = np.where(ellipsoid_mask(shape, ctr, axes) & body, val, mu)
    return act.astype(np.float32), mu.astype(np.float32)


def cameras(n_views, first=0.0, span=2 * np.pi):
    """Angles about the axial axis, exclusive endpoint."""
    return (first + span * np.arange(n_views) / n_views).astype(np.float32)


def workload(name, seed=0):
    """Inputs of a named workload.

    spect64:  64^3, 64 views, PSF 7 (sigma 0.5 -> 2.0), mu 0.15 in the body, 1e6 counts
    spect128: 128^3, 120 views, PSF 11 (0.6 -> 3.0), mu 0.15 + lung-like 0.05, 1e7 counts,
              8 ordered subsets
    spect256: 256^3, 180 views, PSF 21 (0.8 -> 4.5), mu 0.15/0.05/0.30, 5e7 counts
    small:    32^3, 16 views, PSF 5, for quick checks
    """
    if name == "spect64":
        shape, V, K, s0, s1, nl, counts = (64, 64, 64), 64, 7, 0.5, 2.0, 8, 1e6
        act, mu = phantom(shape, nl, seed, body_axes=(28, 24, 28))
    elif name == "spect128":
        shape, V, K, s0, s1, nl, counts = (128, 128, 128), 120, 11, 0.6, 3.0, 12, 1e7
        act, mu = phantom(shape, nl, seed, body_axes=(56, 48, 56), lesion_axes=(4, 16),
                          extra_mu=[((0.35, 0.1, 0.0), (16, 24, 20), 0.05)])
    elif name == "spect256":
        shape, V, K, s0, s1, nl, counts = (256, 256, 256), 180, 21, 0.8, 4.5, 24, 5e7
        act, mu = phantom(shape, nl, seed, body_axes=(112, 96, 112), lesion_axes=(8, 32),
                          extra_mu=[((0.35, 0.1, 0.0), (32, 48, 40), 0.05),
                                    ((-0.4, -0.2, 0.3), (12, 40, 12), 0.30)])
    elif name == "small":
        shape, V, K, s0, s1, nl, counts = (32, 24, 32), 16, 5, 0.5, 1.5, 3, 1e5
        act, mu = phantom(shape, nl, seed, body_axes=(13, 10, 13), lesion_axes=(1.5, 4))
    else:
        raise ValueError(name)
    psf = gaussian_psf(shape[0], K, s0, s1)
    return {"activity": act, "attenuation": mu, "psf": psf, "cameras": cameras(V), "counts": counts,
            "subsets": 8 if name == "spect128" else 1}


def poisson_data(noiseless, counts, seed=0):
    """Scale a noiseless projection to `counts` total and draw Poisson counts."""
    rng = np.random.default_rng(seed)
    lam = noiseless.astype(np.float64) * (counts / max(float(noiseless.sum()), 1e-30))
    return rng.poisson(lam).astype(np.float32)


# Volume shapes (nz, ny, nx) of the batched-reconstruction workload.
BATCH_SHAPES = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 96, 160)]
BATCH_PSF = {7: (0.5, 2.0), 11: (0.6, 3.0), 15: (0.7, 3.8)}


def batch_problem(i):
    """Problem i of the batched workload: seeded shape from BATCH_SHAPES, view
    count U{60..180}, PSF size from {7, 11, 15} with a linear sigma ramp,
    ellipsoid phantom with attenuation 0.15, counts proportional to the volume
    (1e6 at 64^3)."""
    rng = np.random.default_rng(1000 + i)
    shape = BATCH_SHAPES[int(rng.integers(len(BATCH_SHAPES)))]
    V = int(rng.integers(60, 181))
    K = int(rng.choice([7, 11, 15]))
    return batch_problem_of(shape, V, K, 1000 + i)


def batch_problem_of(shape, V, K, seed):
    """A problem of the batched workload's recipe with a given volume shape,
    view count and PSF size."""
    n_les = max(4, shape[2] // 16)
    act, mu = phantom(shape, n_les, seed, lesion_axes=(2.0, max(4.0, shape[2] / 16.0)))
    s0, s1 = BATCH_PSF[K]
    return {"activity": act, "attenuation": mu, "psf": gaussian_psf(shape[0], K, s0, s1), "cameras": cameras(V),
            "counts": 1e6 * act.size / 64 ** 3}
