"""Product host logic (geometry, neighbour list, H0 and Gamma assembly) through
the C ABI, plus the library's exported symbols.  CPU only: none of these calls
touches a device."""
import os
import re

import numpy as np
import pytest

import largetb_amd as L
import oracle as O
from conftest import ROOT, dense_from_triplets


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ltb.h")).read()
    names = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ltb_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 20
    lib = L.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(L.SIGNATURES), set(L.SIGNATURES) ^ names


@pytest.mark.parametrize("Lc,name", [(1, "Amat"), (4, "H")])
def test_product_h_matches_reference_dump(ref_dumps, Lc, name):
    s = L.System.quartz(Lc, Lc, Lc, lattice=L.LATTICE_REFERENCE)
    R = dense_from_triplets(ref_dumps, name)
    assert np.abs(s.dense_h() - R).max() < 2e-15


def test_product_gamma_matches_reference_dump(ref_dumps):
    s = L.System.quartz(4, 4, 4, lattice=L.LATTICE_REFERENCE)
    rp, col, val = s.gamma()
    G = np.zeros((s.nat, s.nat))
    for i in range(s.nat):
        G[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    R = dense_from_triplets(ref_dumps, "GAMMA_AB")
    assert np.abs(G - R).max() < 1e-15
    assert np.abs(G @ ref_dumps["netCharge"] - ref_dumps["GammaQ"]).max() < 4e-15  # |gq| ~ 5


@pytest.mark.parametrize("cells,sigma,seed", [((2, 2, 2), 0.05, 1), ((3, 2, 4), 0.15, 7),
                                              ((5, 5, 4), 0.10, 11)])
def test_product_assembly_matches_oracle(cells, sigma, seed):
    s = L.System.quartz(*cells, sigma_ang=sigma, seed=seed)
    o = O.System.quartz(*cells, sigma_ang=sigma, seed=seed)
    Z, xyz = s.coords()
    assert np.array_equal(Z, o.Z) and np.array_equal(xyz, o.xyz)
    assert s.nnzb == o.nnzb
    assert np.abs(s.dense_h() - o.dense_h()).max() < 5e-15
    rp, col, val = s.hamiltonian()
    for i in range(s.nat):  # diagonal block first, then ascending columns
        assert col[rp[i]] == i
        assert (np.diff(col[rp[i] + 1:rp[i + 1]]) > 0).all()
    assert s.nelectrons == o.nelectrons
    assert np.array_equal(s.neutral(), o.p0)


def test_cell_list_equals_brute_force_scan_on_larger_cell():
    # 6x6x6 uses the cell list; the oracle scans every image of every atom
    s = L.System.quartz(6, 6, 6, sigma_ang=0.12, seed=5)
    o = O.System.quartz(6, 6, 6, sigma_ang=0.12, seed=5)
    rp, col, _ = s.hamiltonian()
    for i in range(0, s.nat, 37):
        mine = sorted(col[rp[i] + 1:rp[i + 1]])
        ref = sorted(o.nbr[i * O.MAXN:i * O.MAXN + o.nn[i]])
        assert mine == ref


def test_failure_same_site_atoms():
    xyz = np.array([[0.0, 0, 0], [0.1, 0, 0], [3.0, 0, 0]])
    with pytest.raises(L.LtbError) as e:
        L.System.from_coords([14, 8, 8], xyz, periodic_mask=0)
    assert e.value.status == L.LTB_ERR_GEOMETRY


def test_failure_too_many_neighbours():
    # a compressed cluster: 22 O atoms within 4 Bohr of a Si
    rng = np.random.default_rng(0)
    pts = [[0.0, 0.0, 0.0]]
    while len(pts) < 23:
        p = rng.uniform(-2.2, 2.2, 3)
        if all(np.linalg.norm(p - q) > 0.5 for q in pts):
            pts.append(p)
    with pytest.raises(L.LtbError) as e:
        L.System.from_coords([14] + [8] * 22, np.array(pts), periodic_mask=0)
    assert e.value.status == L.LTB_ERR_NEIGHBORS


def test_failure_unknown_element_and_nan():
    with pytest.raises(L.LtbError) as e:
        L.System.from_coords([14, 6], np.array([[0.0, 0, 0], [3.0, 0, 0]]), periodic_mask=0)
    assert e.value.status == L.LTB_ERR_GEOMETRY
    with pytest.raises(L.LtbError) as e:
        L.System.from_coords([14, 8], np.array([[0.0, 0, 0], [np.nan, 0, 0]]), periodic_mask=0)
    assert e.value.status == L.LTB_ERR_NONFINITE


def test_context_without_gpu_fails_loudly():
    try:
        c = L.Context(0)
    except L.LtbError as e:
        assert e.status == L.LTB_ERR_DEVICE
        return
    c.close()
    pytest.skip("a GPU is present")
