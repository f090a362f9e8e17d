"""A problem read back from the reference's file layout runs through the
device Newton solve exactly like the problem it was written from (the
imported mesh has no vertex lists: the solve needs dgnodes, bf, t2f, elcon
and the element columns of f only)."""
import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd import binio  # noqa: E402
from hdg_amd.problems import ns2d_square, with_bdf1  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


def test_solve_from_file(ctx, tmp_path):
    from hdg_amd.solve import HdgState, hdg_solve
    P = with_bdf1(ns2d_square(3, 2, seed=71, av_frac=0.3), 0.02, seed=2)
    path = tmp_path / "ns.bin"
    binio.write_file(path, binio.export_problem(P.master, P.app, P.mesh, P.udg, P.UH))
    m, app, mesh, UDG, UH = binio.import_problem(binio.read_file(path))
    print("[gpu] Newton on a problem read from a file", flush=True)
    A = HdgState(ctx, P.master, P.app, P.mesh, P.udg, P.UH, P.sh)
    B = HdgState(ctx, m, app, mesh, UDG, UH, P.sh)
    sa = hdg_solve(A, newton_tol=1e-10, gmres_tol=1e-12)
    sb = hdg_solve(B, newton_tol=1e-10, gmres_tol=1e-12)
    assert sa.residuals == sb.residuals
    ua, ha = A.host()
    ub, hb = B.host()
    assert np.array_equal(ua, ub) and np.array_equal(ha, hb)
    assert sa.residuals[-1] <= 1e-10 * sa.residuals[0]
