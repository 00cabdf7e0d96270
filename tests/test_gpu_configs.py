"""GPU parity at the workloads' real sizes, through the C-ABI, against the FP64 oracle:

- 1024^2 image, oversample 2, KB 6 x 6, M = 2e6: complex64 and complex128 Phi and Phi^dagger, the adjoint
  dot-product test on complex Gaussian x and y, and the 50-iteration power-method norm in FP64;
- 2048^2 image, oversample 2, KB 6 x 6, M = 4e6 (the measured pair): both precisions and the dot test;
- 64 channels (sides 256-1024, M log-uniform in [1e4, 1e6], uv sigma U[0.5, 1.2], KB 4 x 4, complex64, 100
  PADMM iterations each, captured as graphs and run together): every channel in the batch equals the
  channel run alone, bit for bit, and so does every channel of the channel batch (one launch per kernel of
  the iteration for each image size); the channel with the fewest visibilities, the one with the most and
  the sparsest one (visibilities per pixel) reconstruct within 0.05 dB SNR of the oracle's FP64 PADMM;
- 20 requests of the serving mix (sides 256-2048, M log-uniform in [1e4, 4e6], 30 % with w-terms, w in
  [0, 200], 1 degree field): each result is the same alone and in the mix, and each request's Phi and
  Phi^dagger agree with the oracle on a 20000-visibility subset (the forward rows of the subset; the
  adjoint of a vector that is zero outside it).

Tolerances (relative L2): complex128 1e-12, complex64 1e-5; PADMM SNR 0.05 dB.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import algorithms, measurementoperator, serving, utilities, wavelets

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def full_size_parity(n, M, J, seed, power_iters=0):
    u, v = make_uv(seed, M, n)
    x = utilities.synth_image(seed, n, n, 20)
    y = utilities.synth_cnormal(seed + 1, M)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, J, J)
    yo, xo = orc.forward(x), orc.adjoint(y)
    xr, yr = utilities.synth_cnormal(seed + 2, n * n), utilities.synth_cnormal(seed + 3, M)
    norms = {}
    for prec in ("c128", "c64"):
        lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, power_iters, 1e-6, "kb", J, J,
                                                         precision=prec)
        norms[prec] = lt.norm
        lt.op.set_norm(1.0)
        assert rel(lt * x, yo) < TOL[prec], prec
        assert rel(lt.adjoint() * y, xo) < TOL[prec], prec
        lhs, rhs = np.vdot(yr, lt * xr), np.vdot(lt.adjoint() * yr, xr)
        assert abs(lhs - rhs) / abs(lhs) < TOL[prec], prec
        lt.op.close()
        torch.cuda.empty_cache()
    return orc, norms


def test_1024_m2e6_both_precisions_dot_test_and_fp64_norm():
    orc, norms = full_size_parity(1024, 2_000_000, 6, 1001, power_iters=50)
    # the power method runs in FP64 for both precisions: the same norm, and the oracle's to 1e-10
    assert norms["c64"] == norms["c128"]
    assert norms["c128"] == pytest.approx(orc.power_method(50, 1e-6), rel=1e-10)


def test_2048_m4e6_both_precisions_and_dot_test():
    full_size_parity(2048, 4_000_000, 6, 1002)


def test_dense_centre_adjoint_with_compensated_tile_sums():
    # the worst case for FP32 accumulation: 4e6 visibilities concentrated at the centre (uv sigma 0.2 rad,
    # about 38 visibilities per cell at the peak, so the central 16 x 16 tiles sum about 10^4 list
    # entries each); the complex64 tile sums are compensated, so the adjoint stays at FP32 rounding
    n, M, J = 2048, 4_000_000, 6
    u, v = make_uv(1003, M, n, sigma=0.2)
    y = utilities.synth_cnormal(1004, M)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, J, J)
    xo = orc.adjoint(y)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", J, J,
                                                     precision="c64")
    info = lt.op.info()
    assert info.tile_list_len > M
    err = rel(lt.adjoint() * y, xo)
    print(f"dense centre adjoint, complex64: rel. error {err:.2e}", flush=True)
    assert err < 1e-6


# ---- 64 channels -----------------------------------------------------------------------------------------

CH_SEED, N_CH, ITERS = 1022, 64, 100


@pytest.fixture(scope="module")
def channels():
    chans, x0s = [], []
    for ch in utilities.synth_channels(CH_SEED, N_CH):
        n = ch["n"]
        lt = measurementoperator.init_degrid_operator_2d(ch["u"], ch["v"], None, None, n, n, 2.0, 50, 1e-6, "kb", 4,
                                                         4, precision="c64")
        y0 = lt * ch["x0"]
        sigma = utilities.SNR_to_standard_deviation(y0, 30.0)
        chans.append(dict(u=ch["u"], v=ch["v"], y=utilities.add_noise(y0, sigma, ch["seed"] + 2), imsizey=n,
                          imsizex=n, sigma=sigma, operator=lt))
        x0s.append(ch["x0"])
    mc = algorithms.MultiChannelPADMM(chans, itermax=ITERS)
    mc.capture()
    res = mc.run(ITERS)
    assert all(r.niters == ITERS for r in res)
    batch = mc.images()
    yield chans, x0s, mc, batch
    mc.close()
    torch.cuda.empty_cache()


def test_64_channels_batch_equals_each_channel_alone(channels):
    chans, x0s, mc, batch = channels
    assert len(batch) == N_CH
    assert {c["imsizex"] for c in chans} == {256, 512, 1024}
    for k, (_, _, sol) in enumerate(mc.items):
        sol.reset()
        sol.iterate(ITERS, use_graph=True)  # alone: nothing else queued
        assert np.array_equal(sol.image(), batch[k]), k


def test_64_channels_channel_batch_equals_the_stream_batch(channels):
    # one launch per kernel of the iteration for each image size: every channel's image bit for bit as
    # the stream batch's (which equals each channel alone, above)
    chans, x0s, mc, batch = channels
    mc.reset()
    res = mc.run_batched(ITERS)
    assert all(r.niters == ITERS for r in res)
    groups, launches = mc.batch_info()
    print(f"channel batch: {groups} groups, {launches} launches per iteration", flush=True)
    assert groups == 3
    for k, img in enumerate(mc.images()):
        assert np.array_equal(img, batch[k]), k


def oracle_padmm_snr(ch, x0, gpu_norm):
    n = ch["imsizex"]
    orc = oracle.Operator(ch["u"], ch["v"], None, None, n, n, 2.0, oracle.KB, 4, 4)
    orc.set_norm(gpu_norm)
    eps = utilities.calculate_l2_radius(ch["y"].size, ch["sigma"])
    xo, _, _, _ = oracle.padmm(orc, [(wavelets.parse_wavelet("DB4"), 4)], ch["y"], eps, ITERS)
    return oracle.snr_db(x0, xo)


def pick(chans, which):
    M = [c["u"].size for c in chans]
    if which == "fewest":
        return int(np.argmin(M))
    if which == "most":
        return int(np.argmax(M))
    return int(np.argmin([m / c["imsizex"] ** 2 for m, c in zip(M, chans)]))  # sparsest


@pytest.mark.parametrize("which", ["fewest", "sparsest", "most"])
def test_64_channels_snr_within_0p05_db_of_the_oracle(channels, which):
    chans, x0s, mc, batch = channels
    k = pick(chans, which)
    snr_g = oracle.snr_db(x0s[k], batch[k])
    snr_o = oracle_padmm_snr(chans[k], x0s[k], chans[k]["operator"].norm)
    print(f"channel {k} ({which}): {chans[k]['imsizex']}^2, M {chans[k]['u'].size}: SNR GPU {snr_g:.3f} dB, "
          f"oracle {snr_o:.3f} dB", flush=True)
    assert abs(snr_g - snr_o) < 0.05, (k, chans[k]["imsizex"], chans[k]["u"].size, snr_g, snr_o)


# ---- serving mix -----------------------------------------------------------------------------------------


def serving_mix():
    return serving.poisson_requests(1005, 20, rate_per_s=20.0, sizes=(256, 512, 1024, 2048), M_range=(1e4, 4e6),
                                    w_fraction=0.3, wmax=200.0)


def oracle_subset_check(req, res, nsub=20000):
    n, J, M = req.imsizex, req.J, req.u.size
    idx = np.sort(np.random.default_rng(req.rid).choice(M, size=min(nsub, M), replace=False))
    one = np.ones(idx.size, np.complex128)
    if req.w_term:  # the operator's rows take w - mean(w) over all of the request's visibilities
        G = oracle.gridding_matrix_w(req.u[idx], req.v[idx], req.w[idx] - req.w.mean(), one, n, n, 2.0, oracle.KB,
                                     J, J, J, J, req.Jw, oracle.W_APPROX, req.cell, req.cell)
    else:
        G = oracle.gridding_matrix(req.u[idx], req.v[idx], one, n, n, 2.0, oracle.KB, J, J, J, J)
    S = oracle.correction2d(2.0, n, n, oracle.KB, J, J)
    Nv, Nu = oracle.grid_dims(n, n, 2.0)
    yo = G.dot(oracle.fft2d(Nv, Nu, oracle.zero_pad(S, n, n, 2.0, req.x), -1))
    assert rel(res.y[idx], yo) < TOL[req.precision], req.rid
    # the adjoint of a vector that is zero outside the subset, on the GPU at full size
    ysub = np.zeros(M, np.complex128)
    ysub[idx] = req.y[idx]
    lt = measurementoperator.init_degrid_operator_2d(req.u, req.v, req.w, None, n, n, 2.0, 0, 1e-4, "kb", J, J,
                                                     req.w_term, req.Jw, req.cell, req.cell, precision=req.precision)
    xo = oracle.zero_pad_adjoint(S, n, n, 2.0, oracle.fft2d(Nv, Nu, G.adjoint().dot(req.y[idx]), +1))
    assert rel(lt.adjoint() * ysub, xo) < TOL[req.precision], req.rid
    lt.op.close()


def test_serving_mix_alone_equals_mixed_and_matches_the_oracle():
    reqs = serving_mix()
    assert sum(r.w_term for r in reqs) >= 3 and max(r.imsizex for r in reqs) == 2048
    assert max(r.u.size for r in reqs) > 1e6
    mixed_srv = serving.Server(96 << 30, max_inflight=4)
    mixed = [f.result(timeout=300) for f in [mixed_srv.submit(r) for r in reqs]]
    mixed_srv.close()
    assert mixed_srv.peak_inflight > 1
    for r, b in zip(serving_mix(), mixed):
        r.arrival = 0.0
        srv = serving.Server(96 << 30, max_inflight=1)
        a = srv.submit(r).result(timeout=300)
        srv.close()
        assert a.rid == b.rid
        assert np.array_equal(a.y, b.y) and np.array_equal(a.x, b.x), r.rid
        oracle_subset_check(r, a)
    torch.cuda.empty_cache()
