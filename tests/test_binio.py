"""The reference's problem-file layouts (hdg_amd/binio.py):

* "cpp" layout against the reference's own reader, kernel/readInputFile.cpp,
  built from its source into oracle/_ref: a file we write goes through its
  readInput + writeOutput unchanged, byte for byte, and readInput splits it
  into exactly our arrays (every field, the header included);
* "matlab" layout (util/writeBinaryFile.m): write/read round trip, and the
  imported problem gives the same element matrices and Schur outputs, bit
  for bit, as the problem it came from (oracle elemental on both);
* the CBSR pattern equals the trace oracle's elcon2entcon restatement, and the
  face list rebuilt from t2f and bf equals the mesh's.
"""
import os
import subprocess

import numpy as np
import pytest

from hdg_amd import binio
from hdg_amd.problems import euler2d_square, euler3d_cube, gather_uh, ns2d_square, with_bdf1
from oracle import trace_oracle as TO

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
CASES = {
    "ns2d_p2": lambda: ns2d_square(2, 2, seed=61, av_frac=0.5),
    "euler2d_p3_bdf1": lambda: with_bdf1(euler2d_square(2, 3, seed=62, alpha=0.3), 0.1, seed=1),
    "euler3d_p2": lambda: euler3d_cube(1, 2, seed=63, alpha=0.3),
}


def _export(P, layout):
    return binio.export_problem(P.master, P.app, P.mesh, P.udg, P.UH, dt=(0.1, 0.2), layout=layout)


def _ref_tool(name):
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip("the reference reader is not built (oracle/_ref needs the reference sources)")
    return path


@pytest.mark.parametrize("case", list(CASES))
def test_cpp_layout_through_the_reference_reader(tmp_path, case):
    tool, dump = _ref_tool("readinput"), _ref_tool("readinput_dump")
    P = CASES[case]()
    pf = _export(P, "cpp")
    x, y = tmp_path / "x.bin", tmp_path / "y.bin"
    binio.write_file(x, pf)
    subprocess.run([tool, str(x), str(y)], check=True, capture_output=True)
    assert x.read_bytes() == y.read_bytes()
    out = tmp_path / "dump"
    out.mkdir()
    subprocess.run([dump, str(x), str(out)], check=True, capture_output=True)
    hdr = np.fromfile(out / "ndims.bin", dtype=np.int32)
    assert np.array_equal(hdr, binio._header_vector(pf.header).astype(np.int32))
    for name, kind, size in binio.LAYOUTS["cpp"]:
        got = np.fromfile(out / f"{name}.bin", dtype=np.int32 if kind == "i" else np.float64)
        want = np.asarray(pf.arrays[name]).ravel(order="F")
        assert got.size == size(pf.header), name
        assert np.array_equal(got, want.astype(got.dtype)), name
    # and our reader reads what we wrote
    back = binio.read_file(x, "cpp")
    for name in pf.arrays:
        assert np.array_equal(back.arrays[name], np.asarray(pf.arrays[name]).ravel(order="F")), name


@pytest.mark.parametrize("case", list(CASES))
def test_matlab_layout_round_trip(tmp_path, oracle, case):
    P = CASES[case]()
    pf = _export(P, "matlab")
    path = tmp_path / "p.bin"
    binio.write_file(path, pf)
    back = binio.read_file(path, "matlab")
    assert back.header == pf.header
    m, app, mesh, UDG, UH = binio.import_problem(back)
    M0 = P.master
    for k in ("shapvt", "shapvg", "shapvgdotshapvl", "shapft", "shapfg", "shapfgdotshapfc", "shapmv", "shapmf",
              "perm", "permgeom", "plocvl", "gwvl"):
        assert np.array_equal(np.asarray(getattr(m, k)), np.asarray(getattr(M0, k))), k
    assert np.array_equal(mesh.t2f, P.mesh.t2f) and np.array_equal(mesh.elcon, P.mesh.elcon)
    assert np.array_equal(mesh.bf, P.mesh.bf) and np.array_equal(mesh.dgnodes, P.mesh.dgnodes)
    assert np.array_equal(mesh.f[:, -2:], P.mesh.f[:, -2:])
    assert np.array_equal(UDG, P.udg) and np.array_equal(UH, P.UH)
    for k in ("physics", "nd", "ncu", "nc", "localsolve", "tdep", "fc_u", "fc_q", "av_col", "bcm"):
        assert getattr(app, k) == getattr(P.app, k), k
    assert np.array_equal(app.param, P.app.param) and np.array_equal(app.bcs, np.asarray(P.app.bcs))
    # the imported problem is the same problem
    a = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    b = oracle.elemental(m, app, mesh.dgnodes, mesh.bf, UDG, gather_uh(UH, mesh.elcon), P.sh)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("case", list(CASES))
def test_block_crs_and_faces(case):
    P = CASES[case]()
    rowpts, colind = binio.block_crs(P.mesh.t2f)
    r2, c2 = TO.entcon(P.mesh.t2f)
    assert np.array_equal(rowpts, r2) and np.array_equal(colind - 1, c2)
    ef = binio.faces_from_t2f(P.mesh.t2f, P.mesh.bf)
    assert np.array_equal(ef, P.mesh.f[:, -2:])


def test_bad_files(tmp_path):
    P = CASES["ns2d_p2"]()
    pf = _export(P, "matlab")
    path = tmp_path / "p.bin"
    binio.write_file(path, pf)
    raw = path.read_bytes()
    (tmp_path / "short.bin").write_bytes(raw[:-8])
    with pytest.raises(ValueError, match="ends inside"):
        binio.read_file(tmp_path / "short.bin")
    (tmp_path / "long.bin").write_bytes(raw + b"\0" * 8)
    with pytest.raises(ValueError, match="after the last array"):
        binio.read_file(tmp_path / "long.bin")
    bad = binio.ProblemFile(dict(pf.header), dict(pf.arrays), "matlab")
    bad.arrays["UDG"] = bad.arrays["UDG"][:, :, :-1]
    with pytest.raises(ValueError, match="UDG"):
        binio.write_file(tmp_path / "bad.bin", bad)
    # the cpp layout does not carry the physics
    cpp = _export(P, "cpp")
    with pytest.raises(ValueError, match="physics"):
        binio.import_problem(cpp)
    m, app, *_ = binio.import_problem(cpp, physics=P.app.physics)
    assert app.physics == P.app.physics
