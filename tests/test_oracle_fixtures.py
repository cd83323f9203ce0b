"""The oracle against the reference's own dumps (H, Gamma, Gamma.dq, deltaH,
orbital->atom, Jackson kernel, Chebyshev nodes).  CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import dense_from_triplets


@pytest.mark.parametrize("L,name", [(1, "Amat"), (4, "H")])
def test_h_matches_reference_dump(ref_dumps, L, name):
    s = O.System.quartz(L, L, L, mode=O.MODE_REFERENCE)
    H = s.dense_h()
    R = dense_from_triplets(ref_dumps, name)
    assert H.shape == R.shape
    assert np.abs(H - R).max() < 2e-15
    # stored pattern: full 4x4 blocks, explicit zeros included
    pat = np.zeros_like(R, bool)
    pat[ref_dumps[name + "_row"], ref_dumps[name + "_col"]] = True
    mine = np.zeros_like(pat)
    for i in range(s.nat):
        for k in range(s.h_rowptr[i], s.h_rowptr[i + 1]):
            j = s.h_col[k]
            mine[4 * i:4 * i + 4, 4 * j:4 * j + 4] = True
    assert (pat == mine).all()


def test_gamma_matches_reference_dump(ref_dumps):
    s = O.System.quartz(4, 4, 4, mode=O.MODE_REFERENCE)
    G = s.dense_gamma()
    R = dense_from_triplets(ref_dumps, "GAMMA_AB")
    assert np.abs(G - R).max() < 1e-15
    assert (np.diag(G) == -1.0).all()


def test_gamma_times_dq_matches_reference_dump(ref_dumps):
    s = O.System.quartz(4, 4, 4, mode=O.MODE_REFERENCE)
    gq = s.dense_gamma() @ ref_dumps["netCharge"]
    assert np.abs(gq - ref_dumps["GammaQ"]).max() < 2e-15


def test_delta_h_is_repeated_gamma_q(ref_dumps):
    D = dense_from_triplets(ref_dumps, "deltaH")
    gq = ref_dumps["GammaQ"]
    assert np.array_equal(np.diag(D), np.repeat(gq, 4))
    off = D - np.diag(np.diag(D))
    assert not off.any()
    # the dump stores full diagonal 4x4 blocks (explicit zeros)
    rows, cols = ref_dumps["deltaH_row"], ref_dumps["deltaH_col"]
    assert (rows // 4 == cols // 4).all() and len(rows) == 16 * len(gq)


def test_orbital_to_atom_projector(ref_dumps):
    P = dense_from_triplets(ref_dumps, "orbit2Atom")
    expect = np.kron(np.eye(576), np.ones((1, 4)))
    assert np.array_equal(P, expect)


def test_jackson_kernel_matches_gnsfile(ref_dumps):
    g = O.jackson(1000)
    # gnsfile.txt is printed with %lf (6 decimals)
    assert np.abs(g - ref_dumps["gns_col1"]).max() <= 5.0e-7 + 1e-12


def test_chebyshev_nodes_match_dosre(ref_dumps):
    x = O.cheb_nodes(2000)
    # dosre.txt prints xk[i] for i = nTilde-1 .. 1
    assert np.abs(x[::-1][:-1] - ref_dumps["dosre_col1"]).max() <= 5.0e-7 + 1e-12
