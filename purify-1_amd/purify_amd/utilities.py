"""vis_params and the synthetic-input generator (cpp/purify/utilities.h:13-55, utilities.cc:63-101,422-458).

The reference's random_sample_density and add_noise reseed from std::random_device on every draw
(utilities.cc:72-75, 443-446), so its simulations are not reproducible.  The generator here is
seeded: std::mt19937_64 with Box-Muller normals, implemented in the HIP library's host code
(pf_synth_*), so the same seed gives the same inputs in C++, Python and the benchmark.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class vis_units(enum.Enum):  # noqa: N801
    lambda_ = 0
    radians = 1
    pixels = 2


@dataclass
class vis_params:  # noqa: N801
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray = None
    vis: np.ndarray = None
    weights: np.ndarray = None
    units: vis_units = vis_units.lambda_
    ra: float = 0.0
    dec: float = 0.0
    average_frequency: float = 0.0

    def __post_init__(self):
        self.u = np.ascontiguousarray(self.u, np.float64)
        self.v = np.ascontiguousarray(self.v, np.float64)
        M = self.u.size
        self.w = np.zeros(M) if self.w is None else np.ascontiguousarray(self.w, np.float64)
        self.vis = np.ones(M, np.complex128) if self.vis is None else np.ascontiguousarray(self.vis, np.complex128)
        self.weights = (np.ones(M, np.complex128) if self.weights is None
                        else np.ascontiguousarray(self.weights, np.complex128))

    def size(self):
        return self.vis.size


def synth_uv(seed, M, sigma):
    """u, v ~ N(0, sigma) radians, clipped to [-pi, pi]."""
    u = np.empty(M)
    v = np.empty(M)
    _lib.check(_lib.lib().pf_synth_uv(C.c_uint64(seed), M, float(sigma), u.ctypes.data, v.ctypes.data))
    return u, v


def synth_image(seed, ny, nx, nsrc=20):
    """Sum of nsrc Gaussian sources (centre U[0,n)^2, amplitude U[0.1,1], sigma U[1,8] px), complex128."""
    img = np.empty(ny * nx, np.complex128)
    _lib.check(_lib.lib().pf_synth_image(C.c_uint64(seed), ny, nx, nsrc, img.ctypes.data))
    return img


def synth_cnormal(seed, n, scale=1.0):
    """Complex normal vector, E|z|^2 = scale^2."""
    z = np.empty(n, np.complex128)
    _lib.check(_lib.lib().pf_synth_cnormal(C.c_uint64(seed), n, float(scale), z.ctypes.data))
    return z


def radians_to_pixels(u_rad, n, os_=2.0):
    """utilities::uv_scale (utilities.cc:258-278): u / (2 pi) * floor(os * n)."""
    return np.asarray(u_rad) / (2 * np.pi) * np.floor(os_ * n)


def SNR_to_standard_deviation(y0, snr=30.0):
    """utilities.cc:422-432: ||y0||_2 / sqrt(M) * 10^(-snr/20)."""
    y0 = np.asarray(y0)
    return np.linalg.norm(y0) / np.sqrt(y0.size) * 10.0 ** (-snr / 20.0)


def calculate_l2_radius(y_size, sigma=0.0, n_sigma=2.0, distribution="chi"):
    """utilities.cc:396-421: the l2-ball radius epsilon."""
    if distribution == "chi^2":
        r = np.sqrt(2 * y_size + n_sigma * np.sqrt(4 * y_size))
        return r if sigma == 0 else r * sigma
    if distribution == "chi":
        alpha = 1.0 / (8 * y_size) - 1.0 / (128 * y_size * y_size)
        mean = np.sqrt(2) * np.sqrt(y_size) * (1 - alpha)
        sd = np.sqrt(2 * y_size * alpha * (2 - alpha))
        return mean + n_sigma * sd if sigma == 0 else (mean + n_sigma * sd) * sigma
    return 1.0


def add_noise(y0, sigma, seed):
    """utilities.cc:434-458 with a seeded generator: re, im ~ N(0, sigma/sqrt(2))."""
    return np.asarray(y0) + synth_cnormal(seed, np.asarray(y0).size, sigma)


def synth_channels(seed, n_channels, sizes=(256, 512, 1024), M_range=(1e4, 1e6), sigma_range=(0.5, 1.2)):
    """The batched multi-channel workload: per channel an image side drawn from `sizes`, M log-uniform over
    M_range, uv Gaussian with sigma uniform over sigma_range (rad, clipped to [-pi, pi]) in pixels of the
    oversample-2 grid, and a 20-source image.  Returns dicts with n, M, uv_sigma, u, v, x0 and the
    channel's own seed (seed + 7 c; +2 is its noise seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n_channels):
        n = int(rng.choice(list(sizes)))
        M = int(np.exp(rng.uniform(np.log(M_range[0]), np.log(M_range[1]))))
        sig = float(rng.uniform(*sigma_range))
        cs = seed + 7 * c
        u, v = synth_uv(cs, M, sig)
        out.append(dict(n=n, M=M, uv_sigma=sig, u=radians_to_pixels(u, n), v=radians_to_pixels(v, n),
                        x0=synth_image(cs + 1, n, n, 20), seed=cs))
    return out


def read_visibility(vis_name, w_term=False):
    """utilities.cc:140-183: a whitespace-separated text file of rows u v [w] re(V) im(V) sigma (nan and inf
    allowed).  Returns vis_params with v negated (the reference's reflection for the image orientation),
    weights 1 / sigma, units lambda (set ``units`` as the caller's data need)."""
    cols = 6 if w_term else 5
    with open(vis_name) as f:
        vals = np.array([float(t) for t in f.read().split()], np.float64)
    if vals.size % cols:
        raise RuntimeError(f"{vis_name}: {vals.size} numbers do not make rows of {cols}")
    a = vals.reshape(-1, cols)
    u, v = a[:, 0], a[:, 1]
    w = a[:, 2] if w_term else np.zeros(a.shape[0])
    re, im, sig = a[:, cols - 3], a[:, cols - 2], a[:, cols - 1]
    return vis_params(u=u, v=-v, w=w, vis=re + 1j * im, weights=(1.0 / sig).astype(np.complex128))


def write_visibility(uv_vis, file_name, w_term=False):
    """utilities.cc:185-202: rows u -v [w] re(V) im(V) 1/re(weight), 13 significant digits."""
    with open(file_name, "w") as f:
        for i in range(uv_vis.u.size):
            parts = [uv_vis.u[i], -uv_vis.v[i]] + ([uv_vis.w[i]] if w_term else [])
            parts += [uv_vis.vis[i].real, uv_vis.vis[i].imag, 1.0 / uv_vis.weights[i].real]
            f.write(" ".join(f"{x:.13g}" for x in parts) + "\n")
