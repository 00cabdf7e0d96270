"""GPU parity of the compact operators (compact_operators.h) against the oracle.

* operators::init_grid_degrid_operator_2d: Phi^dagger Phi / ||Phi||^2.  The GPU applies G then G^H
  exactly; the reference's explicit G^H G drops entries |value| <= 1e-7 (compact_operators.h:109-111),
  so against that form the bar is the reference's own test bar (tests/operators.cc:171-197).
* operators::init_psf_convolve_2d: psf = Phi^dagger Phi delta, circular convolution by unitary n x n
  FFTs, divided by the square of its own power-method norm (tests/operators.cc:199-215).
Tolerances: complex128 1e-12, complex64 1e-5 relative L2 against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import KIND, make_uv, rel
from purify_amd import measurementoperator, operators, utilities
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("J", [4, 6])
def test_grid_degrid_operator_vs_oracle(precision, J):
    n, M = 64, 5000
    u, v = make_uv(401 + J, M, n)
    wt = utilities.synth_cnormal(402, M)
    gd = operators.init_grid_degrid_operator_2d(u, v, None, wt, n, n, 2.0, 30, 1e-6, "kb", J, J,
                                                precision=precision)
    nrm = gd.norm
    orc = oracle.Operator(u, v, None, wt, n, n, 2.0, oracle.KB, J, J)
    orc.set_norm(nrm)
    x = utilities.synth_image(403, n, n, 20)
    out = gd(x)
    assert rel(out, orc.adjoint(orc.forward(x))) < TOL[precision]
    # against the reference's form: explicit G^H G pruned at 1e-7 (the reference's test bar is 1e-4)
    pruned = oracle.grid_degrid(u, v, wt, n, n, 2.0, x, oracle.KB, J, J, norm=nrm, prune=1e-7)
    assert rel(out, pruned) < 1e-4
    # output-first call form of sopt::OperatorFunction
    buf = np.empty(n * n, np.complex128)
    gd(buf, x)
    assert np.array_equal(buf, out)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_grid_degrid_operator_is_self_adjoint_psd_and_normalised(precision):
    n, M = 64, 8000
    u, v = make_uv(411, M, n)
    gd = operators.init_grid_degrid_operator_2d(u, v, None, None, n, n, 2.0, 200, 1e-8, "kb", 4, 4,
                                                precision=precision)
    x = utilities.synth_cnormal(412, n * n)
    y = utilities.synth_cnormal(413, n * n)
    a, b = np.vdot(y, gd(x)), np.vdot(gd(y), x)
    assert abs(a - b) / abs(a) < TOL[precision]
    assert np.vdot(x, gd(x)).real > 0
    # tests/operators.cc:182-188: the normalised operator's power-method norm is 1
    lam = oracle.power_method_fn(gd, utilities.synth_cnormal(414, n * n), 300, 1e-6)
    assert abs(lam - 1.0) < 1e-3


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_grid_degrid_operator_device_tensors_in_place(precision):
    n, M = 32, 2000
    u, v = make_uv(421, M, n)
    gd = operators.init_grid_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                precision=precision)
    x = utilities.synth_image(422, n, n, 10)
    ref = gd(x)
    dt = torch.complex64 if precision == "c64" else torch.complex128
    xd = torch.from_numpy(x).to(dt).cuda()
    out = gd(xd)
    gd(xd, xd)  # in place
    torch.cuda.synchronize()
    assert torch.equal(out, xd)
    assert rel(out.cpu().numpy(), ref) < (1e-6 if precision == "c64" else 1e-15)


def test_grid_degrid_operator_w_projection_and_given_matrix():
    n, M = 64, 3000
    u, v = make_uv(431, M, n)
    w = np.random.default_rng(432).uniform(0, 150, M)
    gd = operators.init_grid_degrid_operator_2d(u, v, w, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4, True, 4,
                                                3600.0 / n, 3600.0 / n, precision="c128")
    orc = oracle.Operator(u, v, w, None, n, n, 2.0, oracle.KB, 4, 4, True, 4, 3600.0 / n, 3600.0 / n)
    x = utilities.synth_cnormal(433, n * n)
    assert rel(gd(x), orc.adjoint(orc.forward(x))) < 1e-12
    # an operator on a given sparse G (init_degrid_operator_2d_from_matrix): Phi^dagger Phi of that G
    import scipy.sparse as sp
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, 4, 4, 4, 4)
    Gs = sp.csr_matrix((G.vals, G.colidx, G.rowptr), shape=(G.rows, G.cols))
    lt = measurementoperator.init_degrid_operator_2d_from_matrix(Gs, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                                 precision="c128")
    got = lt.op.normal(x)
    assert rel(got, lt.adjoint() * (lt * x)) < 1e-13


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_psf_convolution_vs_oracle(precision):
    n, M = 64, 6000
    u, v = make_uv(441, M, n)
    wt = utilities.synth_cnormal(442, M)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, wt, n, n, 2.0, 40, 1e-6, "kb", 4, 4,
                                                     precision=precision)
    start = utilities.synth_cnormal(443, n * n)
    psf = operators.init_psf_convolve_2d(lt, n, n, 60, 1e-6, start=start)
    orc = oracle.Operator(u, v, None, wt, n, n, 2.0, oracle.KB, 4, 4)
    orc.set_norm(lt.norm)
    # the PSF norm: the same power iteration on the oracle's operator from the same start
    lam = oracle.power_method_fn(lambda z: oracle.psf_convolve(orc, z), start, 60, 1e-6)
    assert psf.norm == pytest.approx(lam, rel=1e-10 if precision == "c128" else 1e-5)
    x = utilities.synth_cnormal(444, n * n)
    assert rel(psf(x), oracle.psf_convolve(orc, x, psf.norm)) < TOL[precision]
    # tests/operators.cc:209-214: the normalised convolution has power-method norm 1
    lam1 = oracle.power_method_fn(psf, start, 100, 1e-6)
    assert abs(lam1 - 1.0) < 1e-3


def test_psf_default_start_and_errors():
    n, M = 32, 1500
    u, v = make_uv(451, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 20, 1e-6, "kb", 4, 4,
                                                     precision="c128")
    psf = operators.init_psf_convolve_2d(lt, n, n, 80, 1e-8)
    # the default pseudo-random start finds the largest eigenvalue, max |FFT(psf)| (the eigenvalues of
    # x -> IFFT(FFT(psf) . FFT(x)) with unitary FFTs)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4)
    orc.set_norm(lt.norm)
    delta = np.zeros(n * n, np.complex128)
    delta[(n // 2 - 1) * n + n // 2 - 1] = 1
    ft = oracle.fft2d(n, n, orc.adjoint(orc.forward(delta)))
    lam_max = np.abs(ft).max()
    assert psf.norm ** 2 == pytest.approx(lam_max, rel=1e-2)
    with pytest.raises(ValueError):
        operators.init_psf_convolve_2d(lt, n, 2 * n)
    with pytest.raises(ValueError):
        operators.init_psf_convolve_2d(lt, start=np.ones(5))
