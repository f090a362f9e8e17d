"""Converts the reference's only golden vector for the elemental path,
master/testpython.mat (the full master struct for a p=2 triangle, pgauss=4,
nodetype 0), into tests/golden/master_p2_tri.npz.  Run once in a container
where the reference tree is mounted; the .npz is committed data."""
import os
import sys

import numpy as np
import scipy.io as sio

SRC = "/root/reference/master/testpython.mat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "master_p2_tri.npz")

FIELDS = ["plocvl", "plocfc", "perm", "gpvl", "gwvl", "gpfc", "gwfc", "shapvl", "shapfc",
          "shapvt", "shapvg", "shapvgdotshapvl", "shapft", "shapfg", "shapfgdotshapfc"]


def main():
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} not found")
    m = sio.loadmat(SRC)["masterm"][0, 0]
    out = {k: np.asarray(m[k], dtype=np.float64) for k in FIELDS}
    for k in ("porder", "pgauss", "nd"):
        out[k] = np.asarray(m[k], dtype=np.int64).ravel()
    np.savez(DST, **out)
    print("wrote", DST)


if __name__ == "__main__":
    main()
