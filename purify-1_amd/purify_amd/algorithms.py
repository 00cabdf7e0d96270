"""Imaging PADMM on the GPU: factory::algorithm_factory / padmm_factory
(cpp/purify/algorithm_factory.h:36-146).

``padmm_factory(dist, measurements, wavelets, uv_data, sigma, imsizey, imsizex, sara_size, ...)`` takes
the reference's arguments and returns an ``ImagingProximalADMM`` whose ``run()`` gives a diagnostic
(``x``, ``niters``, ``good``).  The l1 proximal is the closed form for a tight frame (one wavelet
basis, as main.cc:121-126 selects for one dictionary entry) or, with ``tight_frame=False`` (SARA), the
iterative proximal under the positivity and real constraints (algorithm_factory.h:60-66), which runs
eagerly (its stopping test reads the objective every step).  SOPT is not in the reference tree: the
iterations restate the published algorithms and their agreement with SOPT's iterates is unpinned.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, utilities
from .operators import LinearTransform, MeasurementOperator, check_device_tensor
from .wavelets import WaveletOperator


@dataclass
class Diagnostic:
    x: np.ndarray            # the image, complex128 on the host
    niters: int
    good: bool               # converged (relative variation and residual tolerance met)
    residual_norm: float
    relative_variation: float


class ImagingProximalADMM:
    def __init__(self, phi: MeasurementOperator, psi: WaveletOperator, y, epsilon, gamma=0.0, nu=1.0,
                 lagrange_update_scale=0.9, relative_variation=1e-3, residual_tolerance=0.0, itermax=500,
                 tight_frame=None, l1_proximal_tolerance=1e-2, l1_proximal_itermax=50, l1_proximal_nu=1.0,
                 positivity_constraint=True, real_constraint=True):
        """tight_frame None: the closed-form proximal iff the dictionary has one basis."""
        self.phi, self.psi = phi, psi  # kept alive: the solver uses them
        p = _lib.pf_padmm_params()
        _lib.check(_lib.lib().pf_padmm_params_init(C.byref(p)))
        p.gamma, p.epsilon, p.nu = float(gamma), float(epsilon), float(nu)
        p.lagrange_update_scale = float(lagrange_update_scale)
        p.relative_variation = float(relative_variation)
        p.residual_tolerance = float(residual_tolerance)
        p.itermax = int(itermax)
        p.tight_frame = -1 if tight_frame is None else int(bool(tight_frame))
        p.l1_tolerance, p.l1_itermax, p.l1_nu = float(l1_proximal_tolerance), int(l1_proximal_itermax), float(l1_proximal_nu)
        p.positivity, p.real_constraint = int(bool(positivity_constraint)), int(bool(real_constraint))
        self.itermax = int(itermax)
        self._h = C.c_void_p()
        if hasattr(y, "is_cuda") and y.is_cuda:
            y = check_device_tensor(y.contiguous(), phi.M, phi._dtype(), None, "y")
            self._y = y
            _lib.check(_lib.lib().pf_padmm_create(phi.handle, psi.handle, C.c_void_p(y.data_ptr()), _lib.PF_DEVICE,
                                                  C.byref(p), C.byref(self._h)))
        else:
            yh = np.ascontiguousarray(np.asarray(y), np.complex128).ravel()
            if yh.size != phi.M:
                raise ValueError(f"{yh.size} visibilities for an operator of {phi.M}")
            _lib.check(_lib.lib().pf_padmm_create(phi.handle, psi.handle, yh.ctypes.data, _lib.PF_HOST, C.byref(p),
                                                  C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().pf_padmm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iterate(self, iters, use_graph=True, check_every=0):
        r = _lib.pf_padmm_result()
        _lib.check(_lib.lib().pf_padmm_run(self._h, int(iters), int(bool(use_graph)), int(check_every), C.byref(r)))
        return r

    def enqueue(self, iters, use_graph=True):
        """Queue iterations on the solver's stream without waiting (see sync)."""
        _lib.check(_lib.lib().pf_padmm_enqueue(self._h, int(iters), int(bool(use_graph))))

    def sync(self):
        r = _lib.pf_padmm_result()
        _lib.check(_lib.lib().pf_padmm_sync(self._h, C.byref(r)))
        return r

    def reset(self):
        _lib.check(_lib.lib().pf_padmm_reset(self._h))

    def image(self):
        out = np.empty(self.phi.imsizey * self.phi.imsizex, np.complex128)
        _lib.check(_lib.lib().pf_padmm_get(self._h, _lib.PF_PADMM_IMAGE, out.ctypes.data, _lib.PF_HOST))
        return out

    def residual(self):
        out = np.empty(self.phi.M, np.complex128)
        _lib.check(_lib.lib().pf_padmm_get(self._h, _lib.PF_PADMM_RESIDUAL, out.ctypes.data, _lib.PF_HOST))
        return out

    def stats(self):
        """Per-iteration (||x - x_prev||^2, ||x||^2, ||r||^2), FP64."""
        cnt = C.c_int32()
        _lib.check(_lib.lib().pf_padmm_stats(self._h, None, 0, C.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 3))
        _lib.check(_lib.lib().pf_padmm_stats(self._h, out.ctypes.data, cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def run(self, use_graph=True, check_every=1):
        """SOPT's operator(): iterate until converged (checked every check_every iterations) or itermax."""
        r = self.iterate(self.itermax, use_graph, check_every)
        return Diagnostic(self.image(), int(r.niters), bool(r.converged), float(r.residual_norm),
                          float(r.relative_variation))


class ImagingForwardBackward(ImagingProximalADMM):
    """sopt::algorithm::ImagingForwardBackward as cpp/example/fb_random_coverage.cc:89-106 configures it, on
    the device: x~ = x - (beta / sigma^2)(Phi^dagger Phi x - Phi^dagger y), x = prox_(mu beta ||Psi^dagger .||_1)(x~)
    with the l1 proximal of the PADMM (closed form for one basis, iterative and constrained otherwise),
    from x = 0; converged on the relative variation alone.  mu <= 0 / beta <= 0 take the example's
    1e-5 N and 0.5 / N.  ``residual()`` is the image-domain residual Phi^dagger Phi x_prev - Phi^dagger y."""

    def __init__(self, phi: MeasurementOperator, psi: WaveletOperator, y, sigma, mu=0.0, beta=0.0,
                 relative_variation=1e-3, itermax=500, tight_frame=None, l1_proximal_tolerance=1e-2,
                 l1_proximal_itermax=50, l1_proximal_nu=1.0, positivity_constraint=True, real_constraint=True):
        self.phi, self.psi = phi, psi
        p = _lib.pf_fb_params()
        _lib.check(_lib.lib().pf_fb_params_init(C.byref(p)))
        p.mu, p.beta, p.sigma = float(mu), float(beta), float(sigma)
        p.relative_variation, p.itermax = float(relative_variation), int(itermax)
        p.tight_frame = -1 if tight_frame is None else int(bool(tight_frame))
        p.l1_tolerance, p.l1_itermax, p.l1_nu = float(l1_proximal_tolerance), int(l1_proximal_itermax), float(l1_proximal_nu)
        p.positivity, p.real_constraint = int(bool(positivity_constraint)), int(bool(real_constraint))
        self.itermax = int(itermax)
        self._h = C.c_void_p()
        if hasattr(y, "is_cuda") and y.is_cuda:
            y = check_device_tensor(y.contiguous(), phi.M, phi._dtype(), None, "y")
            self._y = y
            _lib.check(_lib.lib().pf_fb_create(phi.handle, psi.handle, C.c_void_p(y.data_ptr()), _lib.PF_DEVICE,
                                               C.byref(p), C.byref(self._h)))
        else:
            yh = np.ascontiguousarray(np.asarray(y), np.complex128).ravel()
            if yh.size != phi.M:
                raise ValueError(f"{yh.size} visibilities for an operator of {phi.M}")
            _lib.check(_lib.lib().pf_fb_create(phi.handle, psi.handle, yh.ctypes.data, _lib.PF_HOST, C.byref(p),
                                               C.byref(self._h)))

    def residual(self):
        out = np.empty(self.phi.imsizey * self.phi.imsizex, np.complex128)
        _lib.check(_lib.lib().pf_padmm_get(self._h, _lib.PF_PADMM_RESIDUAL, out.ctypes.data, _lib.PF_HOST))
        return out


class algo_distribution:  # noqa: N801 -- algorithm_factory.h:23
    serial = "serial"
    mpi_serial = "mpi_serial"
    mpi_distributed = "mpi_distributed"


def padmm_factory(dist, measurements, wavelets, uv_data, sigma, imsizey, imsizex, sara_size, max_iterations=500,
                  real_constraint=True, positive_constraint=True, tight_frame=False, relative_variation=1e-3,
                  l1_proximal_tolerance=1e-2, maximum_proximal_iterations=50):
    """padmm_factory (algorithm_factory.h:36-122): epsilon = calculate_l2_radius(M, sigma), gamma =
    max |Psi^dagger Phi^dagger y| * 1e-3, nu 1, lagrange_update_scale 0.9, residual tolerance epsilon."""
    if sara_size > 1 and tight_frame:
        raise RuntimeError("l1 proximal not consistent: You say you are using a tight frame, but you have more than "
                           "one wavelet basis.")
    if dist != algo_distribution.serial:
        raise RuntimeError("Type of distributed proximal ADMM algorithm not recognised. You might not have compiled "
                           "with MPI.")
    phi = measurements.op if isinstance(measurements, LinearTransform) else measurements
    eps = utilities.calculate_l2_radius(uv_data.vis.size, sigma)
    return ImagingProximalADMM(phi, wavelets, uv_data.vis, eps, 0.0, 1.0, 0.9, relative_variation, eps,
                               max_iterations, tight_frame=tight_frame, l1_proximal_tolerance=l1_proximal_tolerance,
                               l1_proximal_itermax=maximum_proximal_iterations, l1_proximal_nu=1.0,
                               positivity_constraint=positive_constraint, real_constraint=real_constraint)


class MultiChannelPADMM:
    """Many independent channels (frequency channels or snapshots with their own uv-coverage, image size
    and visibilities) reconstructed together: every channel's operator, dictionary and solver stay
    resident.  Two ways to run them:

      * run_batched: a channel batch (pf_padmm_batch) launches every kernel of the iteration once per
        group of alike channels (same image size and configuration), each group on its own stream, the
        whole lock-step iteration captured as one graph;
      * run: every solver's iteration captured as its own graph, all channels' graphs queued at once on
        the solvers' streams.

    Both give each channel exactly its result alone.  channels: list of dicts with u, v (pixels), y,
    imsizey, imsizex, sigma and optional J (4), levels (4), precision ("c64"), power_iters (50), or with
    "operator": a LinearTransform already built (and normalised) for the channel, used in place of
    building one.  tile: the operators' tile shape, (tile_u, tile_v), or "auto": per class of alike
    channels (image size, support, precision; a class is what the channel batch launches together) the
    shape is chosen by the timed builds of init_degrid_operator_2d(tile="auto") on the class's channel
    with the most visibilities, and every channel of the class is built with it (tile_choice holds the
    choices)."""

    def __init__(self, channels, itermax=100, tile=(0, 0)):
        from .operators import measurementoperator

        def key(ch):
            return (int(ch["imsizey"]), int(ch["imsizex"]), int(ch.get("J", 4)), ch.get("precision", "c64"))

        def build(ch, t):
            ny, nx, J, prec = key(ch)
            return measurementoperator.init_degrid_operator_2d(ch["u"], ch["v"], None, ch.get("weights"), ny, nx, 2.0,
                                                               int(ch.get("power_iters", 50)), 1e-4, "kb", J, J,
                                                               precision=prec, tile=t)

        self.tile_choice = {}
        tuned = {}
        if isinstance(tile, str):
            if tile != "auto":
                raise ValueError('tile must be (tile_u, tile_v) or "auto"')
            rep = {}
            for i, ch in enumerate(channels):
                if ch.get("operator") is None:
                    k = key(ch)
                    if k not in rep or len(ch["u"]) > len(channels[rep[k]]["u"]):
                        rep[k] = i
            for k, i in rep.items():
                tuned[i] = build(channels[i], "auto")
                self.tile_choice[k] = (tuned[i].op.get_tuning("tile_u"), tuned[i].op.get_tuning("tile_v"))
        self.items = []
        for i, ch in enumerate(channels):
            ny, nx, J, prec = key(ch)
            lt = ch.get("operator")
            if lt is None:
                lt = tuned.get(i)
            if lt is None:
                lt = build(ch, self.tile_choice.get(key(ch), tile if not isinstance(tile, str) else (0, 0)))
            lv = int(ch.get("levels", 4))
            while lv > 0 and (ny % (1 << lv) or nx % (1 << lv)):
                lv -= 1
            W = WaveletOperator([("DB4", lv)], ny, nx, prec)
            eps = utilities.calculate_l2_radius(np.asarray(ch["y"]).size, float(ch["sigma"]))
            sol = ImagingProximalADMM(lt.op, W, ch["y"], eps, itermax=itermax)
            self.items.append((lt, W, sol))
        self._batch = C.c_void_p()

    def capture(self):
        """Capture every channel's iteration (one graph each), in this thread: one warm-up iteration is
        run and undone per channel (see pf_padmm_run)."""
        for _, _, sol in self.items:
            sol.enqueue(0, use_graph=True)
            sol.sync()

    def reset(self):
        for _, _, sol in self.items:
            sol.reset()

    def run(self, iters, use_graph=True):
        for _, _, sol in self.items:
            sol.enqueue(iters, use_graph)
        return [sol.sync() for _, _, sol in self.items]

    def _make_batch(self):
        if not self._batch:
            hs = (C.c_void_p * len(self.items))(*[sol._h.value for _, _, sol in self.items])
            _lib.check(_lib.lib().pf_padmm_batch_create(hs, len(self.items), C.byref(self._batch)))
        return self._batch

    def run_batched(self, iters, use_graph=True):
        """iters lock-step iterations of every channel through the channel batch (see the class docstring)."""
        _lib.check(_lib.lib().pf_padmm_batch_run(self._make_batch(), int(iters), int(bool(use_graph))))
        return [sol.sync() for _, _, sol in self.items]

    def batch_info(self):
        """(groups, launches per batched iteration) of the channel batch."""
        g, n = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().pf_padmm_batch_info(self._make_batch(), C.byref(g), C.byref(n)))
        return g.value, n.value

    def images(self):
        return [sol.image() for _, _, sol in self.items]

    def close(self):
        if self._batch:
            _lib.lib().pf_padmm_batch_destroy(self._batch)
            self._batch = C.c_void_p()
        for lt, W, sol in self.items:
            sol.close()
            W.close()
            lt.op.close()
        self.items = []
