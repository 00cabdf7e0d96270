"""GPU parity of the wavelet dictionary against the oracle's restatement (complex128 1e-12, complex64
1e-5 relative L2), the reference test's SARA (tests/wavelet_factory.cc:11-36: Dirac, DB1..DB8 at 3
levels on 128 x 128) included; adjointness and perfect reconstruction on the GPU."""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import rel
from purify_amd import utilities, wavelets
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}
SARA9 = [("Dirac", 3)] + [(f"DB{k}", 3) for k in range(1, 9)]


def entries(ws):
    return [(wavelets.parse_wavelet(n), lv) for n, lv in ws]


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("ws,ny,nx", [([("DB4", 3)], 64, 64), (SARA9, 128, 128), ([("DB2", 2), ("Dirac", 1)], 32, 64),
                                      ([("DB8", 3), ("DB10", 2)], 16, 16), ([("DB1", 5), ("DB2", 4)], 32, 32),
                                      ([("DB3", 3), ("DB6", 2)], 16, 64)])
def test_analysis_and_synthesis_vs_oracle(precision, ws, ny, nx):
    W = wavelets.wavelet_operator_factory("serial", ws, ny, nx, precision=precision)
    x = utilities.synth_cnormal(501, ny * nx)
    c = W.adjoint() * x
    assert rel(c, oracle.sara_analysis(x, ny, nx, entries(ws))) < TOL[precision]
    cc = utilities.synth_cnormal(502, len(ws) * ny * nx)
    assert rel(W * cc, oracle.sara_synthesis(cc, ny, nx, entries(ws))) < TOL[precision]
    # adjoint pair and tight frame: Psi Psi^dagger = I
    assert abs(np.vdot(cc, c) - np.vdot(W * cc, x)) / abs(np.vdot(cc, c)) < TOL[precision]
    assert rel(W * c, x) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_device_tensors(precision):
    ny = nx = 64
    W = wavelets.WaveletOperator([("DB4", 4)], ny, nx, precision)
    dt = torch.complex64 if precision == "c64" else torch.complex128
    x = utilities.synth_image(511, ny, nx, 10)
    xd = torch.from_numpy(x).to(dt).cuda()
    c = W.analysis(xd)
    back = W.synthesis(c)
    torch.cuda.synchronize()
    assert rel(back.cpu().numpy(), x) < TOL[precision]
    assert rel(c.cpu().numpy(), W.adjoint() * x) < (1e-6 if precision == "c64" else 1e-15)


def test_errors():
    with pytest.raises(PurifyError, match="divisible"):
        wavelets.WaveletOperator([("DB4", 3)], 20, 20)
    W = wavelets.WaveletOperator([("DB4", 1)], 8, 8)
    with pytest.raises(ValueError):
        W * np.zeros(10)
