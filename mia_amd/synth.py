"""Seeded synthetic inputs of the benchmark configurations (numpy / scipy).

No network and no datasets: every image is generated here from a seed.
  * smoothed_noise: N(0,1) white noise, Gaussian smoothing (mirror edges,
    truncated at 3 sigma), min/max rescaled to [0, 1].
  * gaussian_blobs: sum of isotropic Gaussian blobs plus N(0, noise) noise.
  * random_spline_params: coefficients N(0, sd) clipped to +-clip per component.
  * spline_warp: the image warped by T(x) = x - u(x) with u the cubic B-spline
    displacement of the coefficients (the reference's convention,
    mia/3d/transform/spline.cc:310-314), cubic B-spline image interpolation
    with mirror boundaries (scipy's order-3 spline with mode 'mirror').
"""
import numpy as np


def smoothed_noise(shape, sigma=4.0, seed=0):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    a = gaussian_filter(rng.standard_normal(size=shape, dtype=np.float32), sigma, mode="mirror", truncate=3.0)
    a = (a - a.min()) / (a.max() - a.min())
    return a.astype(np.float32)


def gaussian_blobs(shape, n_blobs=8, sigma_range=(6.0, 14.0), amp_range=(0.3, 1.0), noise=0.01, seed=0):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    img = np.zeros(shape)
    for _ in range(n_blobs):
        c = [rng.uniform(0, n) for n in shape]
        s = rng.uniform(*sigma_range)
        a = rng.uniform(*amp_range)
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        img += a * np.exp(-r2 / (2 * s * s))
    if noise > 0:
        img += rng.normal(0.0, noise, size=shape)
    return img.astype(np.float32)


def random_spline_params(csize, sd=1.5, clip=4.0, seed=0):
    """csize in (x, y[, z]) order; returns interleaved float64 parameters."""
    rng = np.random.default_rng(seed)
    dim = len(csize)
    p = rng.normal(0.0, sd, size=dim * int(np.prod(csize)))
    return np.clip(p, -clip, clip)


def _axis_matrix(n, cs):
    """Dense n x cs matrix of the forward spline weights along one axis
    (float grid scale and position, 16.16 floor, last sample trimmed to three
    weights), mia/3d/transform/spline.cc:237-273."""
    scale = np.float32(cs - 3) / np.float32(n - 1)
    B = np.zeros((n, cs))
    for i in range(n):
        pos = np.float32(np.float32(i) * scale) + np.float32(1.0)
        f = int(np.floor(np.rint(np.float64(pos) * 65536.0) / 65536.0))
        x = np.float64(pos) - f
        xm1 = 1.0 - x
        w3 = x ** 3 / 6.0
        w0 = xm1 ** 3 / 6.0
        w = [w0, xm1 - 2 * w0 + w3, x - 2 * w3 + w0, w3]
        for k in range(4 if i < n - 1 else 3):
            if f - 1 + k < cs:
                B[i, f - 1 + k] += w[k]
    return B


def spline_displacement(shape, csize, params):
    """u(x) on the voxel grid: shape (numpy order) + (dim,)."""
    dim = len(shape)
    C = np.asarray(params, dtype=np.float64).reshape(tuple(csize[::-1]) + (dim,))
    mats = [_axis_matrix(n, cs) for n, cs in zip(shape, csize[::-1])]
    if dim == 2:
        return np.einsum("yj,xi,jic->yxc", mats[0], mats[1], C)
    t = np.einsum("zk,kjic->zjic", mats[0], C)
    t = np.einsum("yj,zjic->zyic", mats[1], t)
    return np.einsum("xi,zyic->zyxc", mats[2], t)


def spline_warp(image, csize, params):
    from scipy.ndimage import map_coordinates
    shape = image.shape
    dim = len(shape)
    u = spline_displacement(shape, csize, params)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    # numpy axis a corresponds to displacement component dim-1-a
    coords = [g - u[..., dim - 1 - a] for a, g in enumerate(grids)]
    out = map_coordinates(image.astype(np.float64), coords, order=3, mode="mirror", prefilter=True)
    return out.astype(np.float32)


def spline_csize(n, rate, dim):
    """Coefficient grid size of `spline:rate=<rate>` (float arithmetic of
    mia/2d/transform/spline.cc:103-116 and mia/3d/transform/spline.cc:130-143)."""
    r = np.float32(rate)
    if dim == 3:
        v = (np.float32(n - 1) + r) / r
    else:
        v = ((np.float32(n) + r) - np.float32(1.0)) / r
    return int(v) + 2


def series_frames_2d(n_frames=64, side=(96, 224), rate=8.0, max_disp=4.0, seed=0):
    """Frames of the batched 2D series: per frame a Gaussian-blob phantom of
    random size (sides uniform in `side`), quantised to uint16 and scaled to
    float, and the same phantom warped by its own B-spline displacement
    (coefficients N(0, 1.5) clipped to +-max_disp).  Returns [(moving, fixed)]."""
    rng = np.random.default_rng(seed)
    frames = []
    for i in range(n_frames):
        ny, nx = (int(v) for v in rng.integers(side[0], side[1] + 1, size=2))
        fixed = scale_to_unit_range(gaussian_blobs((ny, nx), 8, (6.0, 14.0), (0.3, 1.0), 0.01,
                                                   seed=int(rng.integers(1 << 30))))
        cs = (spline_csize(nx, rate, 2), spline_csize(ny, rate, 2))
        p = random_spline_params(cs, 1.5, max_disp, seed=int(rng.integers(1 << 30)))
        moving = scale_to_unit_range(spline_warp(fixed, cs, p))
        frames.append((moving, fixed))
    return frames


def blob_pair_2d(shape=(128, 128), rate=16.0, sd=1.5, clip=3.0, seed=1):
    """Plain 2D registration input: the fixed image is 8 Gaussian blobs
    (centres uniform in the image, sigma U[6, 14] px, amplitude U[0.3, 1.0])
    plus N(0, 0.01) noise; the moving image is the fixed one warped by a cubic
    B-spline displacement of knot spacing `rate` (coefficients N(0, sd)
    clipped to +-clip px).  Returns (moving, fixed, coefficient grid (x, y),
    true coefficients)."""
    fixed = gaussian_blobs(shape, 8, (6.0, 14.0), (0.3, 1.0), 0.01, seed=seed)
    cs = (spline_csize(shape[1], rate, 2), spline_csize(shape[0], rate, 2))
    p = random_spline_params(cs, sd, clip, seed=seed + 1)
    return spline_warp(fixed, cs, p), fixed, cs, p


def scale_to_unit_range(img, bits=16):
    """uint16 intensities scaled to float32 (config of the batched 2D series)."""
    q = np.round(np.clip(img, 0, None) / max(float(img.max()), 1e-12) * (2 ** bits - 1)).astype(np.uint16)
    return q.astype(np.float32) / np.float32(2 ** bits - 1)
