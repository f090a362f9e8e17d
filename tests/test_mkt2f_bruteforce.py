"""Pins the face numbering of hdg_amd/mesh.py's mkt2f against a literal
restatement of the reference's loop (mesh/mkmesh/mkt2f.m:58-80 with
mesh/mkmesh/mkt2t.m's neighbour table), run by brute force on tiny meshes:

  for i = 1:nt, for j = 1:nfv
    if t2t(i,j) > i or t2t(i,j) <= 0      (neighbour later, or none)
      jf = jf + 1; f(jf,:) = [t(i,face(j,:)), i, ie]; t2f(i,j) = jf
      if ie > 0: t2f(ie, k) = jf for the local face k of ie with those vertices

Indices here are 0-based (a missing neighbour is -1).  Integer work, so the
comparison is exact."""
import numpy as np
import pytest

from hdg_amd.mesh import TET_FACES, TRI_FACES, cubemesh, jitter, mkt2f, squaremesh


def _t2t_bruteforce(t, faces):
    """t2t(i, j): the other element holding every vertex of local face j of i
    (mkt2t.m), -1 on the boundary; O(ne^2) set comparisons."""
    ne, nfe = t.shape[0], faces.shape[0]
    sets = [[frozenset(t[i, faces[j]]) for j in range(nfe)] for i in range(ne)]
    verts = [set(t[i]) for i in range(ne)]
    t2t = -np.ones((ne, nfe), dtype=np.int64)
    for i in range(ne):
        for j in range(nfe):
            for k in range(ne):
                if k != i and sets[i][j] <= verts[k]:
                    t2t[i, j] = k
    return t2t


def mkt2f_loop(t, nd):
    """The loop of mkt2f.m:58-80, statement by statement."""
    faces = TRI_FACES if nd == 2 else TET_FACES
    ne, nfe = t.shape[0], faces.shape[0]
    t2t = _t2t_bruteforce(t, faces)
    f = []
    t2f = -np.ones((ne, nfe), dtype=np.int64)
    for i in range(ne):
        for j in range(nfe):
            ie = t2t[i, j]
            if ie > i or ie < 0:
                jf = len(f)
                f.append(list(t[i, faces[j]]) + [i, ie])
                t2f[i, j] = jf
                if ie >= 0:
                    a = np.sort(t[ie][faces], axis=1)
                    b = np.sort(t[i, faces[j]])
                    k = np.nonzero(np.abs(a - b).sum(axis=1) == 0)[0]
                    t2f[ie, k] = jf
    return np.array(f, dtype=np.int64), t2f


def _shuffled(t, seed):
    """Elements in a random order and each element's vertices rotated, so that
    first visits do not follow the mesh generator's order."""
    rng = np.random.default_rng(seed)
    t = t[rng.permutation(t.shape[0])]
    out = t.copy()
    for e in range(t.shape[0]):
        r = rng.integers(t.shape[1])
        out[e] = np.roll(t[e], r)
    return out


MESHES = {
    "square_3x3": lambda: squaremesh(4, 4)[1],
    "square_5x2_parity1": lambda: squaremesh(6, 3, parity=1)[1],
    "square_4x4_shuffled": lambda: _shuffled(squaremesh(5, 5)[1], 1),
    "square_single": lambda: squaremesh(2, 2)[1][:1],
    "cube_1": lambda: cubemesh(2, 2, 2)[1],
    "cube_2_shuffled": lambda: _shuffled(cubemesh(3, 3, 3)[1], 2),
    "cube_2x1x2": lambda: cubemesh(3, 2, 3)[1],
}


@pytest.mark.parametrize("name", list(MESHES))
def test_mkt2f_matches_reference_loop(name):
    t = MESHES[name]()
    nd = t.shape[1] - 1
    f_ref, t2f_ref = mkt2f_loop(t, nd)
    f, t2f = mkt2f(t, nd)
    assert np.array_equal(f, f_ref)
    assert np.array_equal(t2f, t2f_ref)


def test_loop_restatement_catches_a_wrong_order():
    """A numbering by sorted vertex key (a plausible shortcut) differs from
    the first-visit order the loop produces."""
    t = _shuffled(squaremesh(4, 4)[1], 3)
    f_ref, _ = mkt2f_loop(t, 2)
    key = np.sort(f_ref[:, :2], axis=1)
    by_key = np.lexsort(key.T[::-1])
    assert not np.array_equal(by_key, np.arange(len(f_ref)))


def test_jittered_vertices_do_not_change_numbering():
    """Numbering is combinatorial: moving vertices leaves f and t2f alone."""
    p, t = squaremesh(5, 5)
    q = jitter(p, 0.25, 0.2, seed=4)
    assert not np.allclose(p, q)
    f_ref, t2f_ref = mkt2f_loop(t, 2)
    f, t2f = mkt2f(t, 2)
    assert np.array_equal(f, f_ref) and np.array_equal(t2f, t2f_ref)
