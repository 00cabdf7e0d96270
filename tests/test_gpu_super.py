"""The super operator over visibility blocks (measurementoperator::init_combine_operators /
init_super_operator, operators.h:513-578) against its blocks applied one by one and against the oracle:
the stacked forward, the summed adjoint, the dot test, and the construction checks."""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import measurementoperator, utilities
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu


def _create(uv, n, precision):
    return measurementoperator.init_degrid_operator_2d(uv.u, uv.v, None, uv.weights, n, n, 2.0, 30, 1e-6, "kb", 4, 4,
                                                       precision=precision)


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_super_operator_equals_its_blocks(precision, tol):
    n, M = 64, 5000
    u, v = make_uv(701, M, n)
    w = utilities.synth_cnormal(702, M)
    uv = utilities.vis_params(u=u, v=v, weights=w, units=utilities.vis_units.pixels)
    sup = measurementoperator.init_super_operator(lambda b: _create(b, n, precision), uv, 3)
    assert sup.M == M and [s for _, s in sup.segments] == [1666, 1666, 1668]
    x = utilities.synth_image(703, n, n, 10)
    y = utilities.synth_cnormal(704, M)
    ys = sup * x
    ref = np.concatenate([t * x for t in sup.transforms])
    assert rel(ys, ref) < tol
    xs = sup.adjoint() * y
    xr = sum(t.adjoint() * y[a:a + b] for t, (a, b) in zip(sup.transforms, sup.segments))
    assert rel(xs, xr) < tol
    # each block against the oracle with its own norm
    for t, (a, b) in zip(sup.transforms, sup.segments):
        orc = oracle.Operator(u[a:a + b], v[a:a + b], None, w[a:a + b], n, n, 2.0, oracle.KB, 4, 4)
        orc.set_norm(t.op.norm)
        assert rel(ys[a:a + b], orc.forward(x)) < max(tol, 1e-11)
    # adjointness
    lhs, rhs = np.vdot(y, ys), np.vdot(xs, x)
    assert abs(lhs - rhs) / abs(lhs) < max(tol, 1e-12)
    # device tensors in, device tensors out
    dt = torch.complex64 if precision == "c64" else torch.complex128
    yd = sup * torch.from_numpy(x).to(dt).cuda()
    assert yd.is_cuda and rel(yd.cpu().numpy(), ys) < 1e-6


def test_super_operator_overloads_and_checks():
    n, M = 32, 900
    u, v = make_uv(711, M, n)
    uv = utilities.vis_params(u=u, v=v, units=utilities.vis_units.pixels)
    blocks = [utilities.vis_params(u=u[:400], v=v[:400], units=utilities.vis_units.pixels),
              utilities.vis_params(u=u[400:], v=v[400:], units=utilities.vis_units.pixels)]
    a = measurementoperator.init_super_operator(lambda b: _create(b, n, "c128"), blocks)
    b = measurementoperator.init_super_operator(lambda b: _create(b, n, "c128"), [(0, 400), (400, 500)], uv)
    x = utilities.synth_image(712, n, n, 5)
    assert np.array_equal(a * x, b * x)
    with pytest.raises(RuntimeError, match="does not match"):
        measurementoperator.init_combine_operators(a.transforms, [(0, 400), (400, 499)])
    other = measurementoperator.init_degrid_operator_2d(u, v, None, None, 2 * n, 2 * n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                        precision="c128")
    with pytest.raises(PurifyError, match="share"):
        measurementoperator.init_combine_operators([a.transforms[0], other])
