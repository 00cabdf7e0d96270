"""Chirp-truncated w-projection operators (w_term = 2: the rows of wproj_utilities::wprojection_matrix
applied to the plain Kaiser-Bessel gridding matrix, cpp/purify/wproj_utilities.cc:132-188), built for the
whole operator on the device from the power series of the chirp's spectrum, against the oracle's dense
restatement (one N_g-point FFT per row).

- 64 x 64 grids: every row (through G applied to random grids) and the operator's Phi / Phi^dagger against
  the oracle, complex128 at 1e-12 and complex64 at 1e-5, several energy fractions and kernel supports;
  rows whose threshold falls within rounding of a magnitude are skipped, as for the matrix builder;
- the 2048^2 image, 4e6 visibilities, KB 6 x 6, w in [0, 200], 1 degree field, energy fractions 0.999:
  five rows against the oracle at the real 4096^2 grid (complex128), the footprint within 24 x 24, and the
  complex64 operator's adjoint dot-product test;
- rows of at most 8 x 8 taps are stored as their kept entries: gridding equals the dense rows' bit for bit
  and degridding to rounding (rows at the band's edge included), and the rows match the oracle's.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import measurementoperator, utilities
from test_gpu_wmat import threshold_margin

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def chirp_op(u, v, w, n, J, cell, e_c, e_w, precision, weights=None):
    return measurementoperator.init_degrid_operator_2d(u, v, w, weights, n, n, 2.0, 0, 1e-4, "kb", J, J, True, 6,
                                                       cell, cell, precision=precision, w_projection="chirp",
                                                       energy_chirp=e_c, energy_wproj=e_w)


def oracle_rows(u, v, w, n, J, cell, e_c, e_w, weights=None):
    M = len(u)
    wt = np.ones(M, np.complex128) if weights is None else weights
    G = oracle.gridding_matrix(u, v, wt, n, n, 2.0, oracle.KB, J, J, J, J)
    dense, _ = oracle.wprojection_matrix_dense(G, 2 * n, 2 * n, w, cell, cell, e_c, e_w)
    return G, dense


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("J,e_c,e_w", [(4, 0.999, 0.999), (6, 0.99, 1.0), (6, 0.999, 0.99)])
def test_rows_and_operator_vs_oracle_on_a_64_grid(precision, J, e_c, e_w):
    n, M, cell = 32, 300, 400.0  # a 7.1 degree chirp field: |w| up to 200 spreads the chirps over cells
    u, v = make_uv(1201 + J, M, n)
    w = np.random.default_rng(1202).uniform(-200.0, 200.0, M)
    wt = utilities.synth_cnormal(1203, M) + 1.0
    lt = chirp_op(u, v, w, n, J, cell, e_c, e_w, precision, wt)
    G, dense = oracle_rows(u, v, w, n, J, cell, e_c, e_w, wt)
    assert lt.op.info().w_term == 1 and lt.op.get_tuning("w_footprint") >= J
    # Each row against the oracle's, through G applied to two random grids.  A row may differ only where
    # a threshold lies within rounding of a magnitude, so that the keep decision depends on the last bit:
    # the chirp's spectrum is symmetric, its mirror entries agree to rounding only, and a bisection that
    # does not converge ends between two such entries (mostly at the lower energy fractions).
    tie = np.array([threshold_margin(G, m, 2 * n, w[m], cell, e_c, e_w) < 1e-9 for m in range(M)])
    dt = torch.complex128 if precision == "c128" else torch.complex64
    ok = np.ones(M, bool)
    for seed in (1, 2):
        g = utilities.synth_cnormal(1210 + seed, (2 * n) ** 2)
        yg = lt.op.apply_G(torch.from_numpy(g).to(dt).cuda()).cpu().numpy()
        yo = dense @ g
        scale = np.abs(dense) @ np.abs(g)
        ok &= np.abs(yg - yo) <= TOL[precision] * scale
    assert np.all(ok | tie), np.nonzero(~ok & ~tie)
    assert (~tie).sum() >= 60 and ok.mean() > 0.5
    # the operator: Phi = G_w F Z S, Phi^dagger its adjoint (rows with ties replaced by the GPU's own)
    S = oracle.correction2d(2.0, n, n, oracle.KB, J, J)
    x = utilities.synth_image(1220, n, n, 5)
    yx = dense @ oracle.fft2d(2 * n, 2 * n, oracle.zero_pad(S, n, n, 2.0, x), -1)
    assert rel((lt * x)[ok], yx[ok]) < TOL[precision]
    y = utilities.synth_cnormal(1221, M) * ok
    xo = oracle.zero_pad_adjoint(S, n, n, 2.0, oracle.fft2d(2 * n, 2 * n, dense.conj().T @ y, +1))
    assert rel(lt.adjoint() * y, xo) < TOL[precision]


def test_chirp_rows_construction_errors():
    u = np.zeros(4)
    from purify_amd._lib import PurifyError
    with pytest.raises(PurifyError, match="energy_chirp"):
        chirp_op(u, u, u, 16, 4, 60.0, 1.0, 1.0, "c64")
    with pytest.raises(PurifyError, match="square grid"):
        measurementoperator.init_degrid_operator_2d(u, u, u, None, 16, 32, 2.0, 0, 1e-4, "kb", 4, 4, True, 6, 60.0,
                                                    60.0, w_projection="chirp")


def test_2048_m4e6_w_0_200_rows_footprint_and_dot_test():
    n, M, J = 2048, 4_000_000, 6
    cell = 3600.0 / n  # 1 degree field of view
    u, v = make_uv(1003, M, n)
    w = np.random.default_rng(1004).uniform(0.0, 200.0, M)
    lt = chirp_op(u, v, w, n, J, cell, 0.999, 0.999, "c128")
    K = lt.op.get_tuning("w_footprint")
    info = lt.op.info()
    print(f"chirp rows at 2048^2 / 4e6: footprint {K} x {K}, {info.nnz_w / M:.1f} kept entries per row", flush=True)
    assert J <= K <= 24 and 0 < info.nnz_w <= M * K * K
    # five rows against the oracle's literal construction at the 4096^2 grid
    rows = np.array([0, 17, 123_456, 2_000_001, M - 1])
    G = oracle.gridding_matrix(u[rows], v[rows], np.ones(rows.size, np.complex128), n, n, 2.0, oracle.KB, J, J, J, J)
    dense, _ = oracle.wprojection_matrix_dense(G, 2 * n, 2 * n, w[rows], cell, cell, 0.999, 0.999)
    g = utilities.synth_cnormal(1005, (2 * n) ** 2)
    yg = lt.op.apply_G(torch.from_numpy(g).cuda()).cpu().numpy()
    ok = [threshold_margin(G, k, 2 * n, w[rows[k]], cell, 0.999, 0.999) > 1e-9 for k in range(rows.size)]
    assert sum(ok) >= 4
    for k in range(rows.size):
        if ok[k]:  # to FP64 rounding of the row's sum
            assert abs(yg[rows[k]] - dense[k] @ g) <= 1e-12 * (np.abs(dense[k]) @ np.abs(g)), k
    del dense, lt
    torch.cuda.empty_cache()
    # the complex64 operator of the workload: its adjoint dot-product test
    l64 = chirp_op(u, v, w, n, J, cell, 0.999, 0.999, "c64")
    assert l64.op.get_tuning("w_footprint") == K
    x = utilities.synth_cnormal(1006, n * n)
    y = utilities.synth_cnormal(1007, M)
    lhs, rhs = np.vdot(y, l64 * x), np.vdot(l64.adjoint() * y, x)
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def grid_entry_counts(op):
    """The entries-per-step settings of kept-entry gridding that the operator accepts (1, 2 or 4)."""
    from purify_amd._lib import PurifyError
    q0 = op.get_tuning("w_grid_entries")
    ok = []
    for q in (1, 2, 4):
        try:
            op.set_tuning("w_grid_entries", q)
            ok.append(q)
        except PurifyError:
            pass
    op.set_tuning("w_grid_entries", q0)
    return ok


def check_grid_entries(lt, y, ref, tol):
    """Kept-entry gridding at every accepted entries-per-step setting: one entry per step gives the dense
    rows' sums bit for bit (same additions, same order), more entries the same sums to rounding (each entry
    slot of a step adds into its own copy of the tile, the copies summed in order); every setting repeats
    bit for bit."""
    q0 = lt.op.get_tuning("w_grid_entries")
    counts = grid_entry_counts(lt.op)
    for q in counts:
        lt.op.set_tuning("w_grid_entries", q)
        a = lt.adjoint() * y
        assert np.array_equal(a, lt.adjoint() * y), q
        if q == 1:
            assert np.array_equal(a, ref)
        else:
            assert rel(a, ref) < tol, q
    lt.op.set_tuning("w_grid_entries", q0)
    return counts


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("J", [4, 6])
def test_kept_entry_rows_against_the_dense_rows(precision, J):
    # a 1 degree field with |w| <= 200 keeps the chirp within one cell (footprint J + 2 <= 8): the rows are
    # stored as their kept entries (keep mask, offset, values and tap bytes); the dense rows, rebuilt from
    # them on request, must give the same gridding bit for bit at one entry per step (the same sums in the
    # same order) and to rounding at more, the same degridding to rounding (its lanes take the kept values
    # in stored order), every application must repeat bit for bit, and both match the oracle
    n, M = 32, 3000
    cell = 3600.0 / n
    u, v = make_uv(1301 + J, M, n)
    # rows at the band's edge (the chirp offsets there cross the centred grid's border) and centre
    u[:4] = utilities.radians_to_pixels(np.array([np.pi * 0.999, -np.pi, 0.0, np.pi * 0.5]), n, 2.0)
    v[:4] = utilities.radians_to_pixels(np.array([-np.pi, np.pi * 0.999, np.pi * 0.999, -np.pi * 0.999]), n, 2.0)
    w = np.random.default_rng(1302).uniform(-200.0, 200.0, M)
    wt = utilities.synth_cnormal(1303, M) + 1.0
    lt = chirp_op(u, v, w, n, J, cell, 0.999, 0.999, precision, wt)
    K = lt.op.get_tuning("w_footprint")
    assert K <= 8 and lt.op.get_tuning("w_rows_packed") == 1
    q0 = lt.op.get_tuning("w_grid_entries")
    tol = {"c128": 1e-14, "c64": 1e-6}[precision]
    dt = torch.complex128 if precision == "c128" else torch.complex64
    x = utilities.synth_image(1304, n, n, 5)
    y = utilities.synth_cnormal(1305, M)
    g = torch.from_numpy(utilities.synth_cnormal(1306, (2 * n) ** 2)).to(dt).cuda()
    yd = torch.from_numpy(y).to(dt).cuda()

    def run():
        return (lt * x, lt.adjoint() * y, lt.op.apply_G(g).cpu(), lt.op.apply_GH(yd).cpu(), lt.op.normal(x))

    packed = run()
    for a, b in zip(packed, run()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    lt.op.set_tuning("w_rows_packed", 0)
    assert lt.op.get_tuning("w_rows_packed") == 0
    dense_run = run()
    names = ("forward", "adjoint", "G", "G^H", "normal")
    for name, a, b in zip(names, packed, dense_run):
        a, b = np.asarray(a), np.asarray(b)
        if name in ("adjoint", "G^H") and q0 == 1:
            assert np.array_equal(a, b), name
        else:
            assert rel(a, b) < tol, name
    lt.op.set_tuning("w_rows_packed", 1)
    for a, b in zip(packed, run()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert q0 in check_grid_entries(lt, y, np.asarray(dense_run[1]), tol)
    # the rows against the oracle's (ties in the keep decision skipped, as above)
    G, dense = oracle_rows(u, v, w, n, J, cell, 0.999, 0.999, wt)
    tie = np.array([threshold_margin(G, m, 2 * n, w[m], cell, 0.999, 0.999) < 1e-9 for m in range(M)])
    gh = g.cpu().numpy().astype(np.complex128)
    yo = dense @ gh
    ok = np.abs(packed[2].numpy() - yo) <= TOL[precision] * (np.abs(dense) @ np.abs(gh))
    assert np.all(ok | tie) and (~tie).sum() > M // 2


def test_kept_entry_rows_tuning_errors():
    from purify_amd._lib import PurifyError
    n, M = 32, 100
    u, v = make_uv(1311, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision="c64")
    assert lt.op.get_tuning("w_rows_packed") == 0
    lt.op.set_tuning("w_rows_packed", 0)
    with pytest.raises(PurifyError, match="kept-entry rows"):
        lt.op.set_tuning("w_rows_packed", 1)
    with pytest.raises(PurifyError, match="0 or 1"):
        lt.op.set_tuning("w_rows_packed", 2)
    assert lt.op.get_tuning("w_grid_entries") == 0
    with pytest.raises(PurifyError, match="1, 2 or 4"):
        lt.op.set_tuning("w_grid_entries", 3)
    with pytest.raises(PurifyError, match="kept-entry rows"):
        lt.op.set_tuning("w_grid_entries", 2)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_kept_entry_degridding_slot_counts(precision):
    # the degridding lanes take the kept values in stored order, as many slots each as the largest row
    # needs over 8 lanes (2, 4 or 8): rows pruned hard keep at most 16 taps, rows at energy 1 every
    # non-zero tap of their 6 x 6 or 8 x 8 footprint; every variant must equal the dense rows to rounding
    # and repeat bit for bit
    n, M = 32, 2000
    cell = 3600.0 / n
    u, v = make_uv(1401, M, n)
    w = np.random.default_rng(1402).uniform(-200.0, 200.0, M)
    wt = utilities.synth_cnormal(1403, M) + 1.0
    dt = torch.complex128 if precision == "c128" else torch.complex64
    tol = {"c128": 1e-14, "c64": 1e-6}[precision]
    x = utilities.synth_image(1405, n, n, 5)
    seen = set()
    for J, e_w in ((4, 0.9), (4, 0.999), (4, 1.0), (6, 0.999), (6, 0.9999), (6, 1.0)):
        lt = chirp_op(u, v, w, n, J, cell, 0.999, e_w, precision, wt)
        assert lt.op.get_tuning("w_footprint") <= 8
        slots = lt.op.get_tuning("w_degrid_slots")
        seen.add(slots)
        g = torch.from_numpy(utilities.synth_cnormal(1404, (2 * n) ** 2)).to(dt).cuda()
        yg, yx = lt.op.apply_G(g).cpu().numpy(), lt * x
        assert np.array_equal(lt.op.apply_G(g).cpu().numpy(), yg)
        lt.op.set_tuning("w_rows_packed", 0)
        assert rel(yg, lt.op.apply_G(g).cpu().numpy()) < tol, (J, e_w, slots)
        assert rel(yx, lt * x) < tol, (J, e_w, slots)
        # gridding: as many entries per step as the largest row allows (64 / entries lanes each)
        y = utilities.synth_cnormal(1406, M)
        ref = lt.adjoint() * y
        lt.op.set_tuning("w_rows_packed", 1)
        counts = check_grid_entries(lt, y, ref, tol)
        assert counts == {2: [1, 2, 4], 4: [1, 2], 8: [1]}[slots], (J, e_w, slots, counts)
    assert seen == {2, 4, 8}, seen


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("M", [1, 3, 65])
def test_kept_entry_rows_with_few_visibilities(precision, M):
    # a handful of rows: one owner tile, partial wave steps and batches, the last row's spare entries
    n = 32
    cell = 3600.0 / n
    u, v = make_uv(1501 + M, M, n)
    w = np.random.default_rng(1502 + M).uniform(-200.0, 200.0, M)
    lt = chirp_op(u, v, w, n, 6, cell, 0.999, 0.999, precision)
    assert lt.op.get_tuning("w_rows_packed") == 1
    x = utilities.synth_image(1503, n, n, 3)
    y = utilities.synth_cnormal(1504, M)
    f = lt * x
    lt.op.set_tuning("w_rows_packed", 0)
    tol = {"c128": 1e-14, "c64": 1e-6}[precision]
    assert rel(f, lt * x) < tol
    ref = lt.adjoint() * y
    lt.op.set_tuning("w_rows_packed", 1)
    assert len(check_grid_entries(lt, y, ref, tol)) > 1


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("tile", [(16, 16), (64, 32), (64, 64)])
def test_kept_entry_rows_on_other_tile_shapes(precision, tile):
    # the kept-entry kernels stage and accumulate whole tiles: other tile shapes (owner tiles, halos,
    # list duplication) give the dense rows' results too
    n, M = 64, 20_000
    cell = 3600.0 / n
    u, v = make_uv(1601, M, n)
    w = np.random.default_rng(1602).uniform(-200.0, 200.0, M)
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, 2.0, 0, 1e-4, "kb", 6, 6, True, 6, cell,
                                                     cell, precision=precision, w_projection="chirp",
                                                     energy_chirp=0.999, energy_wproj=0.999, tile=tile)
    assert (lt.op.get_tuning("tile_u"), lt.op.get_tuning("tile_v")) == tile
    assert lt.op.get_tuning("w_rows_packed") == 1
    x = utilities.synth_image(1603, n, n, 5)
    y = utilities.synth_cnormal(1604, M)
    f = lt * x
    lt.op.set_tuning("w_rows_packed", 0)
    tol = {"c128": 1e-14, "c64": 1e-6}[precision]
    assert rel(f, lt * x) < tol
    ref = lt.adjoint() * y
    lt.op.set_tuning("w_rows_packed", 1)
    assert len(check_grid_entries(lt, y, ref, tol)) > 1


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_kept_entry_gridding_with_split_tiles(precision):
    # visibilities crowded into the centre: the densest tiles' lists are cut into parts summed in a fixed
    # order afterwards, by the kept-entry kernels as by the dense one
    n, M = 32, 30_000
    cell = 3600.0 / n
    u, v = make_uv(1701, M, n, sigma=0.3)
    w = np.random.default_rng(1702).uniform(-200.0, 200.0, M)
    lt = chirp_op(u, v, w, n, 6, cell, 0.999, 0.999, precision)
    assert lt.op.get_tuning("split_tiles") > 0
    y = utilities.synth_cnormal(1703, M)
    lt.op.set_tuning("w_rows_packed", 0)
    ref = lt.adjoint() * y
    lt.op.set_tuning("w_rows_packed", 1)
    assert len(check_grid_entries(lt, y, ref, {"c128": 1e-14, "c64": 1e-6}[precision])) > 1


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_autotune_chooses_the_kept_entry_gridding_entries(precision):
    # rows pruned to at most 16 kept taps allow 1, 2 or 4 gridding entries per step: the autotuner times
    # each and keeps one of them, with the same adjoint to rounding, repeating bit for bit
    n, M = 64, 20_000
    cell = 3600.0 / n
    u, v = make_uv(1801, M, n)
    w = np.random.default_rng(1802).uniform(-200.0, 200.0, M)
    lt = chirp_op(u, v, w, n, 4, cell, 0.999, 0.9, precision)
    assert lt.op.get_tuning("w_degrid_slots") == 2
    counts = grid_entry_counts(lt.op)
    assert counts == [1, 2, 4]
    y = utilities.synth_cnormal(1803, M)
    lt.op.set_tuning("w_grid_entries", 1)
    ref = lt.adjoint() * y
    lt.op.autotune(2)
    q = lt.op.get_tuning("w_grid_entries")
    assert q in counts
    a = lt.adjoint() * y
    assert np.array_equal(a, lt.adjoint() * y)
    assert rel(a, ref) < {"c128": 1e-14, "c64": 1e-6}[precision]
