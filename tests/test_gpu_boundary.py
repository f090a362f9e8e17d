"""GPU parity for the boundary models beyond the far field: jittered squares
whose four sides (y = 0, x = 1, y = 1, x = 0) carry different models of the
reference's fbou (application/FM/euler/fbou.m, application/FM/ns/fbou.m), and
the 2D Euler far field with eps > 0 (euler/getan.m:97-103), against the oracle
(pinned by tests/test_oracle_boundary.py) on perturbed states.  FP64 bar 1e-12
normwise per array, as the other parity tests."""
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import perturb, relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.app import euler_app  # noqa: E402
from hdg_amd.problems import euler2d_square, ns2d_square, with_bdf1  # noqa: E402

GAM = 1.4


def _euler_sides(P, minf):
    """bottom wall (2), right pressure from uinf(4) (4), top pressure uinf(1)
    (3), left inflow (5)."""
    pinf = 1.0 / (GAM * minf ** 2)
    ui = np.array(P.app.bcs[0], dtype=np.float64)
    top = ui.copy()
    top[0] = 1.05 * pinf
    P.app.bcm = [2, 4, 3, 5]
    P.app.bcs = np.stack([ui, ui, top, ui])
    return P


def _ns_sides(P):
    """bottom adiabatic wall (2), right pressure (4), top isothermal wall at
    1.1 times the free-stream temperature (3), left Dirichlet at the free
    stream (6)."""
    ui = np.array(P.app.bcs[0], dtype=np.float64)
    wall = np.array([0.0, 0.0, 0.0, 1.1])
    P.app.bcm = [2, 4, 3, 6]
    P.app.bcs = np.stack([ui, ui, wall, ui])
    return P


def _eps(P, eps):
    app = euler_app(2, minf=0.5, eps=eps)
    P.app.param = app.param
    return P


CASES = {
    "euler2d_p2_sides": lambda: _euler_sides(euler2d_square(4, 2, seed=51, minf=0.5), 0.5),
    "euler2d_p3_sides": lambda: _euler_sides(euler2d_square(3, 3, seed=52, minf=0.5, alpha=0.3), 0.5),
    "euler2d_p4_sides": lambda: _euler_sides(euler2d_square(3, 4, seed=53, minf=0.5), 0.5),
    "euler2d_p2_eps": lambda: _eps(euler2d_square(4, 2, seed=54, minf=0.5), 0.4),
    "ns2d_p2_sides": lambda: _ns_sides(ns2d_square(3, 2, seed=55, av_frac=0.3)),
    "ns2d_p3_sides": lambda: _ns_sides(ns2d_square(4, 3, seed=56, av_frac=0.3)),
    "ns2d_p3_sides_bdf1": lambda: with_bdf1(_ns_sides(ns2d_square(3, 3, seed=57, av_frac=0.3)), 0.05, seed=5),
}


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


@pytest.mark.parametrize("case", list(CASES))
def test_boundary_models_element_matrices(ctx, oracle, case):
    P = perturb(CASES[case](), 6)
    print(f"[gpu] boundary models, element matrices {case} ne={P.ne} bcm={list(P.app.bcm)}", flush=True)
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    ref = oracle.elemental_matrices(*args)
    got = ctx.elemental_matrices(*args)
    ctx.stream.synchronize()
    for k, v in ref.items():
        err = relerr(got[k].numpy(), v)
        assert err < 1e-12, (k, err)
    # every side's model is on some element face
    assert set(np.unique(P.mesh.bf[P.mesh.bf < 0])) == {-1, -2, -3, -4}


@pytest.mark.parametrize("case", list(CASES))
def test_boundary_models_elemental(ctx, oracle, case):
    P = perturb(CASES[case](), 7)
    print(f"[gpu] boundary models, elemental {case} ne={P.ne}", flush=True)
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    ref = oracle.elemental(*args)
    got = [o.numpy() for o in ctx.elemental(*args)]
    ctx.stream.synchronize()
    assert np.all(got[4] == 0)
    for name, g, r in zip(("AE", "FE", "DUDG", "DUDG_DUH"), got[:4], ref[:4]):
        assert relerr(g, r) < 1e-12, (name, relerr(g, r))


def test_unsupported_model_is_refused(ctx):
    from hdg_amd.runtime import HdgxError
    P = euler2d_square(2, 2, seed=58)
    P.app.bcm = [1, 1, 6, 1]            # NS's Dirichlet model: not a 2D Euler case
    with pytest.raises(HdgxError):
        ctx.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
