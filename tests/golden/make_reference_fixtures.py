"""Convert the reference's own MATLAB/ASCII dumps into compact numpy fixtures.

Run in a container where the reference tree is mounted (it is read as data only;
no reference code is executed).  Output: ``tests/golden/reference_dumps.npz``.

Sources (all written by PETSc's ASCII-MATLAB viewer or by fprintf in
withSP2_deltaH.cpp):

* ``Amat.m``          H of one alpha-quartz cell, reference lattice (36x36 AIJ,
                      explicit zeros kept) -- assembled by constructH
                      (withSP2_deltaH.cpp:969-1066).
* ``H.m``             H of the 4x4x4 supercell (2304x2304, 32768 stored
                      entries = 2048 full 4x4 blocks), dump code at
                      withSP2_deltaH.cpp:1504-1509.
* ``GAMMA_AB.m``      Gamma matrix (576x576), withSP2_deltaH.cpp:950-955.
* ``orbit2Atom.m``    orbital->atom projector (576x2304), :958-963.
* ``netCharge.m``     delta q per atom (576), :1472-1477.
* ``GammaQ_vector.m`` Gamma . delta q (576), :1479-1484.
* ``deltaH.m``        the diagonal Coulomb shift as 4x4 blocks, :1465-1470.
* ``gnsfile.txt``     column 1 = Jackson kernel g_n for 1000 moments
                      (:1593-1595, dump :1649-1652).
* ``dosre.txt``       column 1 = Chebyshev nodes x_k (withSP2.cpp:1304-1311).
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_dumps.npz")


def read_petsc_mat(path):
    size = None
    rows, cols, vals = [], [], []
    with open(path) as f:
        for line in f:
            m = re.match(r"% Size = (\d+) (\d+)", line)
            if m:
                size = (int(m.group(1)), int(m.group(2)))
                continue
            parts = line.split()
            if len(parts) == 3 and parts[0].isdigit() and parts[1].isdigit():
                rows.append(int(parts[0]) - 1)
                cols.append(int(parts[1]) - 1)
                vals.append(float(parts[2]))
    return size, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals, np.float64)


def read_petsc_vec(path):
    vals = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith(("%", "V", "]")):
                continue
            vals.append(float(s))
    return np.array(vals, np.float64)


def read_cols(path, col):
    out = []
    with open(path) as f:
        for line in f:
            parts = line.replace(",", " ").split()
            if len(parts) > col:
                out.append(float(parts[col]))
    return np.array(out, np.float64)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not mounted; fixtures are already committed")
    d = {}
    for name in ["Amat", "H", "GAMMA_AB", "orbit2Atom", "deltaH"]:
        size, r, c, v = read_petsc_mat(os.path.join(REF, name + ".m"))
        d[name + "_shape"] = np.array(size, np.int32)
        d[name + "_row"] = r
        d[name + "_col"] = c
        d[name + "_val"] = v
    d["netCharge"] = read_petsc_vec(os.path.join(REF, "netCharge.m"))
    d["GammaQ"] = read_petsc_vec(os.path.join(REF, "GammaQ_vector.m"))
    d["gns_col1"] = read_cols(os.path.join(REF, "gnsfile.txt"), 0)
    d["dosre_col1"] = read_cols(os.path.join(REF, "dosre.txt"), 0)
    np.savez_compressed(OUT, **d)
    for k, v in d.items():
        print(k, v.shape, v.dtype)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
