#!/usr/bin/env python
"""Throughput of the measurement-operator pair Phi x, Phi^dagger y on one MI355X.

Workload (default): 2048 x 2048 image, oversample 2 (4096 x 4096 grid), Kaiser-Bessel 6 x 6,
M = 4e6 visibilities with (u, v) ~ N(0, pi/3) rad clipped to [-pi, pi], no w-term, unit weights;
synthetic inputs from the library's seeded generator (seed 1002).  A step is one Phi x followed by
one Phi^dagger y on inputs resident in HBM.  value = M * steps / elapsed (visibilities per second)
for the complex64 operator; the complex128 operator is timed on the same inputs for the FP32/FP64
speed ratio.  Accuracy is the relative L2 error against the FP64 oracle on the same inputs.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--workload pair_2048_m4e6|pair_1024_m2e6|pair_256_m1e4|wproj_2048_m4e6]

Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "purify-1_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOADS = {
    # name: (image side, J, M, seed)
    "pair_2048_m4e6": (2048, 6, 4_000_000, 1002),
    "pair_1024_m2e6": (1024, 6, 2_000_000, 1001),
    "pair_256_m1e4": (256, 4, 10_000, 1000),
    # w-projection rows (w_term): w ~ U[0, 200] wavelengths, 1 degree field of view, Jw = 6
    "wproj_2048_m4e6": (2048, 6, 4_000_000, 1003),
    # chirp-truncated w-projection rows (wprojection_matrix of the plain G, energy fractions 0.999), w ~ U[0, 200]
    # wavelengths, 1 degree field of view; 200 PADMM iterations beside the pair
    "wchirp_2048_m4e6": (2048, 6, 4_000_000, 1003),
}
WPROJ = {"wproj_2048_m4e6": dict(Jw=6, wmax=200.0, fov_deg=1.0),
         "wchirp_2048_m4e6": dict(Jw=6, wmax=200.0, fov_deg=1.0, chirp=0.999)}

METRIC = "Φ+Φ† vis/s FP32 vs FP64 (rel.err≤1e-5, frac HBM peak); served p50/p99 latency"


def row_bytes(KK, nnz, M, e, packed):
    """Stored bytes per projection row: KK taps dense, or a 16-byte record (keep mask, offset) and the kept
    entries with one tap byte each."""
    return 16 + (e + 1) * nnz / M if packed else KK * e


def model_bytes_pair(n, Ng, M, e, r, RB=None):
    """Stage-boundary compulsory traffic of one Phi + Phi^dagger pair.  With stored projection rows of RB
    bytes (w_term) each direction reads per visibility its row origin (8 B), its value and its row."""
    if RB:
        return 3 * e * n * n + 11 * e * Ng + 2 * (8 + e + RB) * M
    return 3 * e * n * n + 11 * e * Ng + (4 * r + 4 * e) * M


def kernel_alg_bytes(name, n, Nu, Nv, M, e, r, RB=None):
    """Algorithmic bytes of one launch of each kernel: its inputs read once and outputs written once."""
    Ng = Nu * Nv
    win = n * Nu  # grid rows that hold the image (forward) or are kept by the crop (adjoint)
    vis = M * (2 * r + 2 * e)  # u, v, weight and value per visibility
    wvis = M * (8 + e + (RB or 0))  # row origin, value and stored projection row
    table = {
        "degrid_wrows": Ng * e + wvis,
        "grid_wrows": wvis + Ng * e,
        "fft_rows_fwd_sz": n * n * e + win * e,
        "fft_rows_adj_crop": win * e + n * n * e,
        "degrid": Ng * e + vis,
        "grid_tiles": vis + Ng * e,
        "fft_cols_stepB": 2 * Ng * e,
        "fft_cols_fwd": (win + Ng) * e,  # reads the image window's rows, writes the full grid
        "fft_cols_adj": (Ng + win) * e,  # reads the full grid, writes the rows the crop keeps
        # the image-column-first order (one-pass column plans): the column pass on the n image columns,
        # the row pass over every grid row
        "fft_cols_fwd_x": (n * n + Nv * n) * e,  # reads the image, writes Nv x n
        "fft_rows_fwd_x": (Nv * n + Ng) * e,  # reads Nv x n, writes the full grid
        "fft_rows_adj_x": (Ng + Nv * n) * e,  # reads the full grid, writes its image columns
        "fft_cols_adj_x": (Nv * n + n * n) * e,  # reads Nv x n, writes the image
        "permute_to_sorted": 2 * M * e,
        "permute_to_caller": 2 * M * e,
    }
    if name == "fft_cols_stepA":
        return None  # depends on direction (window read forward, full read adjoint): averaged below
    return table.get(name)


def serving_run(base, log, precision="c64"):
    """Served latency: a Poisson stream (20 requests / s) of imaging jobs, image side uniform over
    {256, 512, 1024, 2048}, M log-uniform in [1e4, 4e6], 30 % with w-terms; each job builds its operator
    and applies one Phi + Phi^dagger pair, or (25 % of the jobs) normalises it and runs a 100-iteration
    DB4 PADMM reconstruction.  Latency runs from arrival to the result on the host, so it includes the
    operator build and the host transfers.  The same requests then arrive all at once: the served rate of
    that burst is the server's capacity (the Poisson leg's rate is bounded by the arrival rate).  `base`:
    the stream's requests; both precisions serve the same ones."""
    import dataclasses
    from purify_amd import serving

    def serve(reqs):
        srv = serving.Server(96 << 30, max_inflight=4)
        t0 = time.perf_counter()
        futs = [srv.submit(r) for r in reqs]
        res = [f.result() for f in futs]
        wall = time.perf_counter() - t0
        srv.close()
        return serving.latency_stats(res), len(res), wall, srv.peak_inflight

    reqs = [dataclasses.replace(r, precision=precision) for r in base]
    M_total = sum(int(r.u.size) for r in reqs)
    span = reqs[-1].arrival if reqs else 0.0
    st, n, wall, peak = serve(reqs)
    log(f"[bench] serving {precision}: p50 {1e3 * st['p50_s']:.1f} ms, p99 {1e3 * st['p99_s']:.1f} ms, "
        f"{n / wall:.1f} req/s over a {span:.1f} s arrival span")
    bst, bn, bwall, _ = serve([dataclasses.replace(r, arrival=1e-9) for r in reqs])  # all pending at the start
    out = {"requests": n, "rate_per_s": 20.0, "p50_ms": 1e3 * st["p50_s"], "p99_ms": 1e3 * st["p99_s"],
           "mean_ms": 1e3 * st["mean_s"], "throughput_req_per_s": n / wall, "vis_pairs_per_s": M_total / wall,
           "capacity_req_per_s": bn / bwall, "burst_p50_ms": 1e3 * bst["p50_s"], "burst_p99_ms": 1e3 * bst["p99_s"],
           "peak_inflight": peak, "hbm_budget_gib": 96, "w_term_jobs": sum(r.w_term for r in reqs),
           "padmm_jobs": sum(r.kind == "padmm" for r in reqs), "precision": precision, "arrival_span_s": span,
           "sample": (f"the full stream: {n} requests at 20 / s" if n >= 2000 else
                      f"{n} requests of the serving mix (the full stream is 2000 at 20 / s)")}
    log(f"[bench] serving {precision}: capacity {out['capacity_req_per_s']:.1f} req/s (all {bn} requests at once)")
    return out


def sol_gamma(op, W, y, eps):
    """The PADMM's gamma for these data (max |Psi^dagger Phi^dagger y| * 1e-3, algorithm_factory.h:70-73)."""
    from purify_amd import algorithms
    s = algorithms.ImagingProximalADMM(op, W, y, eps, itermax=1)
    r = s.sync()
    s.close()
    return r.gamma


def padmm_run(iters, log, with_cpu=True):
    """PADMM iterations / s with a DB4 (4 levels) tight-frame prior, 30 dB noise, KB, oversample 2:
    the plain-method shape (256^2 image, M = 1e4, J = 4; the oracle's FP64 PADMM on the host beside it,
    and the reconstruction SNR of both against the true image) and the large shape (2048^2,
    M = 4e6, J = 6).  Each leg builds the operator (normalised by its power method), then times `iters`
    iterations replayed from a captured hipGraph and the same number launched eagerly; in complex64 also
    forward-backward (graph) and, at 256^2, the SARA PADMM (nine bases, iterative proximal, eager)."""
    import numpy as np
    import torch
    from purify_amd import algorithms, measurementoperator, utilities, wavelets
    out = {}
    for name, n, M, J in (("plain_256", 256, 10_000, 4), ("pair_2048_m4e6", 2048, 4_000_000, 6)):
        seed = 3001 + n
        u, v = utilities.synth_uv(seed, M, np.pi / 3)
        u, v = utilities.radians_to_pixels(u, n), utilities.radians_to_pixels(v, n)
        x0 = utilities.synth_image(seed + 1, n, n, 20)
        leg = {"image": n, "M": M, "J": J, "wavelet": "DB4", "levels": 4, "iterations": iters}
        for prec in ("c64", "c128"):
            lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 50, 1e-6, "kb", J, J,
                                                             precision=prec)
            if prec == "c64":
                y0 = lt * x0
                sigma = utilities.SNR_to_standard_deviation(y0, 30.0)
                y = utilities.add_noise(y0, sigma, seed + 2)
                eps = utilities.calculate_l2_radius(M, sigma)
            W = wavelets.WaveletOperator([("DB4", 4)], n, n, prec)
            rates = {}
            for graph in (True, False):
                sol = algorithms.ImagingProximalADMM(lt.op, W, y, eps, itermax=iters + 1)
                sol.iterate(1, use_graph=graph)  # capture (graph) / warm-up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                sol.iterate(iters, use_graph=graph)
                torch.cuda.synchronize()
                rates["graph" if graph else "eager"] = iters / (time.perf_counter() - t0)
                if graph:
                    leg[f"{prec}_snr_db"] = float(20 * np.log10(np.linalg.norm(x0) / np.linalg.norm(x0 - sol.image())))
                del sol
            leg[f"{prec}_it_per_s"] = rates
            log(f"[bench] padmm {name} {prec}: {rates['graph']:.0f} it/s graph, {rates['eager']:.0f} it/s eager, "
                f"SNR {leg[prec + '_snr_db']:.2f} dB")
            if prec == "c64":
                # forward-backward (PhiTPhi form, DB4, graph replay): lambda = mu sigma^2 equal to the
                # PADMM's gamma, unit step
                gam = float(sol_gamma(lt.op, W, y, eps))
                fb = algorithms.ImagingForwardBackward(lt.op, W, y, sigma, mu=gam / sigma ** 2, beta=sigma ** 2,
                                                       itermax=iters + 1)
                fb.iterate(1, use_graph=True)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fb.iterate(iters, use_graph=True)
                torch.cuda.synchronize()
                leg["fb_c64_it_per_s"] = iters / (time.perf_counter() - t0)
                leg["fb_c64_snr_db"] = float(20 * np.log10(np.linalg.norm(x0) / np.linalg.norm(x0 - fb.image())))
                del fb
                log(f"[bench] forward-backward {name} c64: {leg['fb_c64_it_per_s']:.0f} it/s graph, "
                    f"SNR {leg['fb_c64_snr_db']:.2f} dB")
                if n <= 512:
                    # SARA (Dirac + DB1..DB8, 3 levels), the iterative constrained proximal (eager)
                    S = wavelets.WaveletOperator([("Dirac", 3)] + [(f"DB{k}", 3) for k in range(1, 9)], n, n, prec)
                    sa = algorithms.ImagingProximalADMM(lt.op, S, y, eps, itermax=iters)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    r = sa.iterate(iters, use_graph=False)
                    leg["sara_c64_it_per_s"] = iters / (time.perf_counter() - t0)
                    leg["sara_c64_l1_steps_per_it"] = r.l1_iters_total / iters
                    leg["sara_c64_snr_db"] = float(20 * np.log10(np.linalg.norm(x0) / np.linalg.norm(x0 - sa.image())))
                    del sa, S
                    log(f"[bench] padmm SARA {name} c64: {leg['sara_c64_it_per_s']:.0f} it/s eager, "
                        f"{leg['sara_c64_l1_steps_per_it']:.1f} proximal steps / it, SNR {leg['sara_c64_snr_db']:.2f} dB")
            del W, lt
            torch.cuda.empty_cache()
        if with_cpu and name == "plain_256":
            import oracle
            orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, J, J, power_iters=50, power_tol=1e-6)
            t0 = time.perf_counter()
            xo, _, _, _ = oracle.padmm(orc, [(4, 4)], y, eps, iters)
            dt = time.perf_counter() - t0
            leg["cpu_oracle"] = {"it_per_s": iters / dt, "snr_db": oracle.snr_db(x0, xo), "cores": oracle.omp_threads(),
                                 "kind": "port"}
            leg["snr_gap_db_c64_vs_oracle"] = leg["c64_snr_db"] - leg["cpu_oracle"]["snr_db"]
            log(f"[bench] padmm {name} oracle (FP64, host): {iters / dt:.1f} it/s, SNR {leg['cpu_oracle']['snr_db']:.2f} dB")
        out[name] = leg
    return out


def multichannel_run(n_channels, iters, seed, log):
    """Batched multi-channel PADMM: n_channels independent channels, image side drawn from
    {256, 512, 1024}, M log-uniform in [1e4, 1e6], uv Gaussian with sigma U[0.5, 1.2] rad/px, KB J = 4,
    oversample 2, complex64 operators normalised by 50 power iterations, DB4 (4 levels) prior, 30 dB
    noise.  Every channel's iteration is captured as a hipGraph; `iters` iterations per channel are
    timed once channel after channel (each waits for its own) and once queued together on the
    channels' streams."""
    import numpy as np
    import torch
    from purify_amd import algorithms, measurementoperator, utilities
    chans, x0s = [], []
    t0 = time.perf_counter()
    specs = list(utilities.synth_channels(seed, n_channels))
    # the tile shape per class of alike channels (here: image side), chosen by the timed builds of
    # tile="auto" on the class's channel with the most visibilities (as MultiChannelPADMM(tile="auto"))
    rep = {}
    for i, ch in enumerate(specs):
        if ch["n"] not in rep or ch["u"].size > specs[rep[ch["n"]]]["u"].size:
            rep[ch["n"]] = i
    tuned, tile_choice = {}, {}
    for n, i in rep.items():
        ch = specs[i]
        tuned[i] = measurementoperator.init_degrid_operator_2d(ch["u"], ch["v"], None, None, n, n, 2.0, 50, 1e-6,
                                                               "kb", 4, 4, precision="c64", tile="auto")
        tile_choice[n] = (tuned[i].op.get_tuning("tile_u"), tuned[i].op.get_tuning("tile_v"))
    for i, ch in enumerate(specs):
        n, u, v, x0 = ch["n"], ch["u"], ch["v"], ch["x0"]
        lt = tuned.get(i)
        if lt is None:
            lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 50, 1e-6, "kb", 4, 4,
                                                             precision="c64", tile=tile_choice[n])
        y0 = lt * x0
        sigma = utilities.SNR_to_standard_deviation(y0, 30.0)
        chans.append(dict(u=u, v=v, y=utilities.add_noise(y0, sigma, ch["seed"] + 2), imsizey=n, imsizex=n,
                          sigma=sigma, operator=lt))
        x0s.append(x0)
    mc = algorithms.MultiChannelPADMM(chans, itermax=iters)
    mc.capture()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    M_total = sum(int(ch["u"].size) for ch in chans)
    t0 = time.perf_counter()
    for _, _, sol in mc.items:
        sol.iterate(iters, use_graph=True)
    t_seq = time.perf_counter() - t0
    mc.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = mc.run(iters)
    t_bat = time.perf_counter() - t0
    imgs = mc.images()
    snr = [float(20 * np.log10(np.linalg.norm(x0) / np.linalg.norm(x0 - x))) for x0, x in zip(x0s, imgs)]
    # the channel batch: recorded, its first iteration eager and one iteration captured, then timed anew
    mc.reset()
    mc.run_batched(2)
    mc.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_cb = mc.run_batched(iters)
    t_cb = time.perf_counter() - t0
    groups, launches = mc.batch_info()
    cb_same = all(np.array_equal(a, b) for a, b in zip(mc.images(), imgs))
    mc.close()
    del chans
    torch.cuda.empty_cache()
    out = {"channels": n_channels, "iterations": iters, "M_total": M_total,
           "image_sides": {str(k): sum(x.size == k * k for x in x0s) for k in (256, 512, 1024)},
           "batched_channel_it_per_s": n_channels * iters / t_bat, "batched_vis_it_per_s": M_total * iters / t_bat,
           "sequential_channel_it_per_s": n_channels * iters / t_seq, "batched_speedup": t_seq / t_bat,
           "batched_s": t_bat, "setup_s": setup, "all_ran": all(r.niters == iters for r in res + res_cb),
           "channel_batch": {"s": t_cb, "channel_it_per_s": n_channels * iters / t_cb,
                             "vis_it_per_s": M_total * iters / t_cb, "speedup_vs_sequential": t_seq / t_cb,
                             "speedup_vs_stream_batch": t_bat / t_cb, "groups": groups,
                             "launches_per_iteration": launches, "equals_stream_batch": cb_same},
           "snr_db": {"min": min(snr), "median": float(np.median(snr)), "max": max(snr)},
           "tile_per_image_side": {str(k): list(t) for k, t in tile_choice.items()}}
    log(f"[bench] multichannel {n_channels} x {iters} it: channel batch {t_cb:.2f} s ({groups} groups, {launches} "
        f"launches / iteration, {n_channels * iters / t_cb:.0f} channel-it/s), stream batch {t_bat:.2f} s "
        f"({out['batched_channel_it_per_s']:.0f} channel-it/s), sequential {t_seq:.2f} s, setup {setup:.1f} s, "
        f"SNR median {out['snr_db']['median']:.2f} dB, channel batch equals stream batch: {cb_same}")
    return out


def wchirp_padmm_run(u, v, w, n, J, cell, chirp_kw, iters, seed, log):
    """PADMM on the chirp-row w-projection operator (complex64, normalised by 50 FP64 power iterations):
    DB4 (4 levels) tight-frame prior, 30 dB noise, `iters` iterations replayed from a captured hipGraph."""
    import numpy as np
    import torch
    from purify_amd import algorithms, measurementoperator, utilities, wavelets
    t0 = time.perf_counter()
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, 2.0, 50, 1e-6, "kb", J, J, True, 6, cell,
                                                     cell, precision="c64", **chirp_kw)
    setup = time.perf_counter() - t0
    x0 = utilities.synth_image(seed + 5, n, n, 20)
    y0 = lt * x0
    sigma = utilities.SNR_to_standard_deviation(y0, 30.0)
    y = utilities.add_noise(y0, sigma, seed + 6)
    eps = utilities.calculate_l2_radius(u.size, sigma)
    W = wavelets.WaveletOperator([("DB4", 4)], n, n, "c64")
    sol = algorithms.ImagingProximalADMM(lt.op, W, y, eps, itermax=iters + 1)
    sol.iterate(1, use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = sol.iterate(iters, use_graph=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    snr = float(20 * np.log10(np.linalg.norm(x0) / np.linalg.norm(x0 - sol.image())))
    out = {"wchirp_2048_m4e6": {"image": n, "M": int(u.size), "J": J, "wavelet": "DB4", "levels": 4,
                                "iterations": iters, "c64_it_per_s": {"graph": iters / dt}, "c64_snr_db": snr,
                                "setup_s_with_fp64_norm": setup, "residual_norm": r.residual_norm,
                                "epsilon": eps}}
    log(f"[bench] padmm wchirp c64: {iters / dt:.0f} it/s graph, SNR {snr:.2f} dB, setup {setup:.1f} s")
    del sol, W, lt
    torch.cuda.empty_cache()
    return out


def counters_run(args, measurementoperator, u, v, x_h, y_h, n, J, os_, dev, w=None, w_term=False, Jw=6, cell=1.0,
                 chirp_kw=None):
    """Workload for rocprofv3 --pmc passes: a copy of known size, then the operator pair."""
    import torch
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)  # 512 MiB
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    prec = args.counters
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, os_, 0, 1e-4, "kb", J, J, w_term, Jw, cell,
                                                     cell, precision=prec, **(chirp_kw or {}))
    op = lt.op
    op.autotune(3)
    dt = torch.complex64 if prec == "c64" else torch.complex128
    x = torch.from_numpy(x_h).to(dt).to(dev)
    yin = torch.from_numpy(y_h).to(dt).to(dev)
    y = torch.empty(len(u), dtype=dt, device=dev)
    xo = torch.empty(n * n, dtype=dt, device=dev)
    for _ in range(args.warmup + args.steps):
        op.forward(x, out=y)
        op.adjoint_apply(yin, out=xo)
    torch.cuda.synchronize()
    print(json.dumps({"counters_run": prec, "pairs": args.warmup + args.steps,
                      "grid_kernel": op.get_tuning("grid_kernel")}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--workload", default="pair_2048_m4e6", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-pairs", type=int, default=2)
    ap.add_argument("--padmm-iters", type=int, default=100, help="PADMM iterations per leg; 0 skips the PADMM legs")
    ap.add_argument("--channels", type=int, default=64, help="channels of the batched multi-channel PADMM leg; 0 skips it")
    ap.add_argument("--serve-requests", type=int, default=400,
                    help="requests of the serving sample (Poisson arrivals at 20 / s); 0 skips it")
    ap.add_argument("--counters", choices=("c64", "c128"), default=None,
                    help="hardware-counter run: three 512 MiB plain copies (the calibration kernel for "
                         "FETCH_SIZE / WRITE_SIZE) then warmup + steps pairs of one operator, nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one operator per GPU, no "
              "collectives); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from purify_amd import measurementoperator, utilities

    n, J, M, seed = WORKLOADS[args.workload]
    os_ = 2.0
    Nu = Nv = int(n * os_)
    Ng = Nu * Nv
    dev = torch.device("cuda:0")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    u_rad, v_rad = utilities.synth_uv(seed, M, np.pi / 3)
    u = utilities.radians_to_pixels(u_rad, n, os_)
    v = utilities.radians_to_pixels(v_rad, n, os_)
    x_h = utilities.synth_image(seed, n, n, 20)
    y_h = utilities.synth_cnormal(seed + 1, M)
    wp = WPROJ.get(args.workload)
    w_term, Jw, cell, w = False, 6, 1.0, None
    if wp:
        w_term, Jw = True, wp["Jw"]
        cell = wp["fov_deg"] * 3600.0 / n  # arcsec per pixel
        w = np.random.default_rng(seed + 2).uniform(0.0, wp["wmax"], M)
    log(f"[bench] workload {args.workload}: n={n} grid={Nu} J={J} M={M} w_term={w_term}")

    chirp_kw = {}
    if wp and wp.get("chirp"):
        chirp_kw = dict(w_projection="chirp", energy_chirp=wp["chirp"], energy_wproj=wp["chirp"])
    if args.counters:
        counters_run(args, measurementoperator, u, v, x_h, y_h, n, J, os_, dev, w, w_term, Jw, cell, chirp_kw)
        return
    res = {}
    outs = {}
    for prec, steps in (("c64", args.steps), ("c128", args.fp64_steps)):
        t0 = time.perf_counter()
        # the product's own tuning: tile shape and permutation bucket by timed builds (tile="auto"), then
        # the gridding kernel and column FFT mode (autotune)
        lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, os_, 0, 1e-4, "kb", J, J, w_term, Jw,
                                                         cell, cell, precision=prec, tile="auto", **chirp_kw)
        setup = time.perf_counter() - t0
        op = lt.op
        tiles = {"tile": [op.get_tuning("tile_u"), op.get_tuning("tile_v")], "perm_bucket": op.get_tuning("perm_bucket")}
        KK_op = op.get_tuning("w_footprint") ** 2 if w_term else None
        packed = op.get_tuning("w_rows_packed") if w_term else 0
        # inputs resident before the autotuner runs, so that the warm-up pairs follow its timed pairs with
        # no host-side gap (the clocks reach their steady state after some tens of milliseconds of load)
        dt = torch.complex64 if prec == "c64" else torch.complex128
        x = torch.from_numpy(x_h).to(dt).to(dev)
        yin = torch.from_numpy(y_h).to(dt).to(dev)
        y = torch.empty(M, dtype=dt, device=dev)
        xo = torch.empty(n * n, dtype=dt, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        variant = op.autotune(3)
        tune_s = time.perf_counter() - t0
        for _ in range(args.warmup):
            op.forward(x, out=y)
            op.adjoint_apply(yin, out=xo)
        torch.cuda.synchronize()
        # per-step latencies with events on torch's current stream (the stream the kernels run on)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record()
        for k in range(steps):
            op.forward(x, out=y)
            op.adjoint_apply(yin, out=xo)
            ev[k + 1].record()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        lat = [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]
        # per-kernel durations (HIP events around each launch, same stream)
        op.profile(True)
        for _ in range(max(3, min(steps, 10))):
            op.forward(x, out=y)
            op.adjoint_apply(yin, out=xo)
        torch.cuda.synchronize()
        table = op.profile_table()
        op.profile(False)
        outs[prec] = (y.cpu().numpy().astype(np.complex128), xo.cpu().numpy().astype(np.complex128))
        # PCIe-inclusive: the same pair on host complex128 buffers (copies and conversions included)
        host_ms = []
        for _ in range(3):
            t0 = time.perf_counter()
            op.forward(x_h)
            op.adjoint_apply(y_h)
            host_ms.append(1e3 * (time.perf_counter() - t0))
        res[prec] = dict(setup_s=setup, steps=steps, elapsed=elapsed, ms=1e3 * elapsed / steps,
                         value=M * steps / elapsed, lat=lat, table=table, grid_kernel=variant, tune_s=tune_s,
                         host_ms=statistics.median(host_ms), KK=KK_op, nnz_w=op.info().nnz_w, packed=packed,
                         fft_columns=op.get_tuning("fft_columns"), fft_row_blocks=op.get_tuning("fft_row_blocks"),
                         grid_entries=op.get_tuning("w_grid_entries") if packed else 0, **tiles)
        log(f"[bench] {prec}: {res[prec]['ms']:.3f} ms/pair, {res[prec]['value']:.3e} vis/s, setup {setup:.2f} s, "
            f"grid kernel {variant}")
        del lt, op, x, yin, y, xo
        torch.cuda.empty_cache()

    r64 = res["c64"]
    e, r = 8, 4
    # dominant kernel of the complex64 pair
    tab = r64["table"]
    dom = max(tab, key=lambda k: tab[k][0])
    dom_ms = tab[dom][0] / tab[dom][1]
    KK = r64["KK"] if w_term else None
    RB = row_bytes(KK, r64["nnz_w"], M, e, r64["packed"]) if KK else None
    RB128 = row_bytes(KK, res["c128"]["nnz_w"], M, 16, res["c128"]["packed"]) if KK else None
    b = kernel_alg_bytes(dom, n, Nu, Nv, M, e, r, RB)
    if b is None:  # step A: forward reads the image window, adjoint the full grid
        b = ((n * Nu * e + Ng * e) + (2 * Ng * e)) / 2
    achieved = b / (dom_ms * 1e-3)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None, "kernel": dom, "kernel_ms": dom_ms,
                "alg_bytes_per_launch": b}
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        per = tr.get("bytes_per_launch", {})
        # the profiles name kernels by their trace entries: the column FFT per direction (mean of the
        # forward and adjoint launches, as the timing averages them), the gridding by its variant
        alias = {"fft_cols_fwd": ["fft_cols_big_fwd"], "fft_cols_adj": ["fft_cols_big_adj"],
                 "grid_tiles": ["grid_tiles[mfma]"],
                 "permute_to_caller": ["permute_to_caller[chunk]", "permute_to_caller[bucket]"],
                 "permute_to_sorted": ["permute_to_sorted[chunk]", "permute_to_sorted[bucket]"]}
        if dom in per:
            roofline["traffic"] = per[dom]
        elif dom in alias and all(k in per for k in alias[dom]):
            vals = [per[k] for k in alias[dom]]
            roofline["traffic"] = sum(vals)
        roofline["traffic_source"] = os.path.relpath(traffic_file, ROOT)

    pair_bytes = model_bytes_pair(n, Ng, M, e, r, RB)
    # the bytes this run moves: unit weights are not stored, so neither direction reads them
    moved_bytes = pair_bytes - (0 if KK else 2 * e * M)
    kern_frac = {}
    for k in tab:
        kb = kernel_alg_bytes(k, n, Nu, Nv, M, e, r, RB)
        if kb:
            kern_frac[k] = kb / (tab[k][0] / tab[k][1] * 1e-3) / HBM_PEAK
    rel_fp32_fp64 = {
        "forward": float(np.linalg.norm(outs["c64"][0] - outs["c128"][0]) / np.linalg.norm(outs["c128"][0])),
        "adjoint": float(np.linalg.norm(outs["c64"][1] - outs["c128"][1]) / np.linalg.norm(outs["c128"][1])),
    }

    cpu = None
    acc_oracle = None
    if not args.no_cpu and not w_term:
        import oracle
        threads = oracle.omp_threads()
        log(f"[bench] cpu baseline: FP64 oracle, explicit CSR G/G^H + OpenMP SpMV, {threads} threads")
        t0 = time.perf_counter()
        orc = oracle.Operator(u, v, None, None, n, n, os_, oracle.KB, J, J)
        build_s = time.perf_counter() - t0
        times = []
        for _ in range(args.cpu_pairs):
            t0 = time.perf_counter()
            yo = orc.forward(x_h)
            xo = orc.adjoint(y_h)
            times.append(time.perf_counter() - t0)
        t_pair = statistics.median(times)
        cpu = {"value": M / t_pair, "unit": "vis/s", "cores": threads, "kind": "port",
               "sample": f"full workload, {args.cpu_pairs} pairs after a {build_s:.1f} s CSR build; "
                         f"median pair {t_pair:.3f} s"}
        acc_oracle = {
            "c64_forward": float(np.linalg.norm(outs["c64"][0] - yo) / np.linalg.norm(yo)),
            "c64_adjoint": float(np.linalg.norm(outs["c64"][1] - xo) / np.linalg.norm(xo)),
            "c128_forward": float(np.linalg.norm(outs["c128"][0] - yo) / np.linalg.norm(yo)),
            "c128_adjoint": float(np.linalg.norm(outs["c128"][1] - xo) / np.linalg.norm(xo)),
        }
        log(f"[bench] cpu: {t_pair:.3f} s/pair; rel err vs oracle {acc_oracle}")
        del orc

    serve = None
    if args.serve_requests > 0 and not w_term:
        from purify_amd import serving
        t0 = time.perf_counter()
        base = serving.poisson_requests(seed + 10, args.serve_requests, 20.0, padmm_fraction=0.25, padmm_iters=100)
        log(f"[bench] serving: {len(base)} requests generated in {time.perf_counter() - t0:.1f} s")
        serve = {p: serving_run(base, log, p) for p in ("c64", "c128")}
        del base
    padmm = None
    if args.padmm_iters > 0 and not w_term:
        padmm = padmm_run(args.padmm_iters, log, with_cpu=not args.no_cpu)
    multi = None
    if args.channels > 0 and args.padmm_iters > 0 and not w_term:
        multi = multichannel_run(args.channels, args.padmm_iters, seed + 20, log)
    if chirp_kw and args.padmm_iters > 0:
        padmm = wchirp_padmm_run(u, v, w, n, J, cell, chirp_kw, 2 * args.padmm_iters, seed, log)

    lat = sorted(r64["lat"])
    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": args.workload, "image": n, "grid": Nu, "oversample": os_, "kernel": "kb",
                   "J": J, "M": M, "w_term": w_term, "uv_sigma_rad": round(np.pi / 3, 6), "seed": seed,
                   **({"Jw": Jw, "w_max_lambda": wp["wmax"], "cell_arcsec": cell} if wp else {}),
                   **({"w_projection": "chirp", "energy_chirp": wp["chirp"], "energy_wproj": wp["chirp"]}
                      if chirp_kw else {})},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pair_frac_hbm_peak": pair_bytes / (r64["ms"] * 1e-3) / HBM_PEAK,
        "pair_model_bytes": pair_bytes,
        "pair_frac_hbm_peak_moved": moved_bytes / (r64["ms"] * 1e-3) / HBM_PEAK,
        "pair_moved_bytes": moved_bytes,
        "kernels_frac_hbm_peak": kern_frac,
        **({"w_rows": {"footprint": int(round(KK ** 0.5)), "kept_entries_per_row": r64["nnz_w"] / M,
                       "layout": "kept entries" if r64["packed"] else "dense",
                       "stored_bytes_per_row": RB,
                       "grid_entries_per_step": {"c64": r64["grid_entries"],
                                                 "c128": res["c128"]["grid_entries"]}}} if KK else {}),
        "fp64": {"value": res["c128"]["value"], "ms_per_step": res["c128"]["ms"],
                 "fp32_speedup": res["c128"]["ms"] / r64["ms"],
                 "pair_frac_hbm_peak": model_bytes_pair(n, Ng, M, 16, 8, RB128) / (res["c128"]["ms"] * 1e-3) / HBM_PEAK},
        "rel_err_fp32_vs_fp64_gpu": rel_fp32_fp64,
        "rel_err_vs_oracle": acc_oracle,
        "step_latency_ms": {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))]},
        "serving": serve,
        "padmm": padmm,
        "multichannel_padmm": multi,
        "setup_s": {"c64": r64["setup_s"], "c128": res["c128"]["setup_s"]},
        "host_pair_ms": {"c64": r64["host_ms"], "c128": res["c128"]["host_ms"],
                         "note": "pair on host complex128 buffers: PCIe copies and conversions included"},
        "autotune": {"grid_kernel": {"c64": r64["grid_kernel"], "c128": res["c128"]["grid_kernel"]},
                     "fft_columns": {"c64": r64["fft_columns"], "c128": res["c128"]["fft_columns"]},
                     "fft_row_blocks": {"c64": r64["fft_row_blocks"], "c128": res["c128"]["fft_row_blocks"]},
                     "tile": {"c64": r64["tile"], "c128": res["c128"]["tile"]},
                     "perm_bucket": {"c64": r64["perm_bucket"], "c128": res["c128"]["perm_bucket"]},
                     "seconds": {"c64": r64["tune_s"], "c128": res["c128"]["tune_s"]},
                     "setup_seconds_with_tile_search": {"c64": r64["setup_s"], "c128": res["c128"]["setup_s"]}},
        "kernels_ms_per_launch": {k: tab[k][0] / tab[k][1] for k in tab},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
