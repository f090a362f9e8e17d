"""GPU parity on the benchmark states themselves, unperturbed: the 2D Euler
p = 2 state at full size (2 x 32 x 32 triangles, free stream M 0.5 with three
1 % Gaussian bumps in rho and p), the Navier-Stokes p = 3 state (free stream
M 0.3, 2 % random-phase sinusoidal velocity, q = -grad u, artificial viscosity
1e-3 on 5 % of the elements) on 2 x 16 x 16 triangles, and the 3D Euler p = 2
state (M 0.4, 1 % density bumps) on 6 x 3^3 tetrahedra.

Bars, per output array, normwise relative to the oracle:
  * AE, FE, DUDG_DUH: 1e-12 in FP64, 1e-10 in FP32 + FP64 refinement.
  * DUDG: these smooth states are nearly discrete equilibria, so the element
    residual (Ru, Rq) that DUDG solves for is a sum of terms up to 1e4 times
    larger than itself, and every FP64 evaluation of it carries a rounding
    error of eps times the terms.  The test measures that floor -- the
    oracle's own change under +-1 ulp perturbations of the state and trace --
    and holds the GPU to max(bar, 4 x floor).  Three more checks close the
    gap the floor leaves open:
      - the elimination alone: the oracle's Schur step run on the GPU's own
        element matrices reproduces the GPU's DUDG to the bar (same inputs);
      - backward error: the GPU's DUDG and DUDG_DUH solve the GPU's element
        systems with a normwise backward error below 1e-12;
      - the element matrices themselves are at 1e-12, and Ru, Rq at the
        floor of the residual.
The reference's arithmetic is FP64 throughout (kernel/schur_uq.c:40-287)."""
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.gpu_util import backward_error, relerr, ulp_perturb

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

from hdg_amd.problems import euler2d_square, euler3d_cube, ns2d_square  # noqa: E402

SEED = 2024
CASES = {
    "euler2d_p2_2x32x32": lambda: euler2d_square(32, 2, seed=SEED),
    "ns2d_p3_2x16x16": lambda: ns2d_square(16, 3, seed=SEED),
    "euler3d_p2_6x3^3": lambda: euler3d_cube(3, 2, seed=SEED),
}
NAMES = ("AE", "FE", "DUDG", "DUDG_DUH")


@pytest.fixture(scope="module")
def ctx():
    from hdg_amd.runtime import Context
    return Context(0)


_cache = {}


def _problem(oracle, case):
    """The problem, the oracle's outputs and element matrices, and the FP64
    floors of DUDG, Ru and Rq (the oracle's response to +-1 ulp in udg, uh)."""
    if case not in _cache:
        P = CASES[case]()
        args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
        ref = oracle.elemental(*args)
        mats = oracle.elemental_matrices(*args)
        rng = np.random.default_rng(99)
        pu, ph = ulp_perturb(P.udg, rng), ulp_perturb(P.uh, rng)
        pert = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, pu, ph, P.sh)
        pm = oracle.elemental_matrices(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, pu, ph, P.sh)
        floor = {"DUDG": relerr(pert[2], ref[2]), "Ru": relerr(pm["Ru"], mats["Ru"])}
        if "Rq" in mats:
            floor["Rq"] = relerr(pm["Rq"], mats["Rq"])
        _cache[case] = (P, ref, mats, floor)
    return _cache[case]


@pytest.mark.parametrize("case", list(CASES))
def test_element_matrices_at_baseline_state(ctx, oracle, case):
    P, ref, mats, floor = _problem(oracle, case)
    print(f"[gpu] baseline-state element matrices {case} ne={P.ne}", flush=True)
    got = ctx.elemental_matrices(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    ctx.stream.synchronize()
    errs = {k: relerr(got[k].numpy(), v) for k, v in mats.items()}
    print(f"[gpu]   errors {errs} floors {floor}", flush=True)
    for k, e in errs.items():
        bar = max(1e-12, 4 * floor[k]) if k in ("Ru", "Rq") else 1e-12
        assert e <= bar, (k, e, bar)
    # the floor is the rounding of a well-posed residual, not a broken state
    assert floor["Ru"] < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("case", list(CASES))
def test_elemental_at_baseline_state(ctx, oracle, case, prec):
    P, ref, mats, floor = _problem(oracle, case)
    bar = 1e-12 if prec == 0 else 1e-10
    print(f"[gpu] baseline-state elemental {case} prec={prec} ne={P.ne}", flush=True)
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    ctx.set_precision(prec)
    try:
        got = [o.numpy() for o in ctx.elemental(*args)]
        gm = {k: v.numpy() for k, v in ctx.elemental_matrices(*args).items()}
        ctx.stream.synchronize()
    finally:
        ctx.set_precision(0)
    assert np.all(got[4] == 0)
    errs = {n: relerr(g, r) for n, g, r in zip(NAMES, got[:4], ref[:4])}
    print(f"[gpu]   vs oracle {errs}; DUDG floor {floor['DUDG']:.2e}", flush=True)
    for n in ("AE", "FE", "DUDG_DUH"):
        assert errs[n] <= bar, (n, errs[n])
    assert errs["DUDG"] <= max(bar, 4 * floor["DUDG"]), (errs["DUDG"], floor["DUDG"])
    assert floor["DUDG"] < 1e-9
    # the elimination on identical inputs: the oracle's Schur step on the
    # GPU's own element matrices
    if P.app.localsolve:
        same = oracle.schur_uq(*(gm[k] for k in ("M", "C", "E", "BD", "F", "GK", "H", "Rq", "Ru", "Rh")))
    else:
        same = oracle.schur_u(*(gm[k] for k in ("BD", "F", "GK", "H", "Ru", "Rh")))
    # the oracle's Schur outputs come as (DUDG, DUDG_DUH, AE, FE)
    e_same = {n: relerr(got[i], same[j]) for n, i, j in (("AE", 0, 2), ("FE", 1, 3), ("DUDG", 2, 0),
                                                         ("DUDG_DUH", 3, 1))}
    print(f"[gpu]   elimination on the GPU's matrices vs the oracle's {e_same}", flush=True)
    for n, e in e_same.items():
        assert e <= bar, (n, e)
    # backward error of the local solves against the GPU's element systems
    eta_u, eta_h = backward_error(P, gm, got[2], got[3])
    print(f"[gpu]   backward error DUDG {eta_u:.2e}, DUDG_DUH {eta_h:.2e}", flush=True)
    assert eta_u < 1e-12 and eta_h < 1e-12
