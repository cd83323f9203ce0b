"""Neighbour list, H0 and Gamma assembled on the device (ltb_load_coords)
against the host assembly (ltb_system_create / ltb_system_quartz, itself
pinned to the reference's H.m / GAMMA_AB.m dumps and to the oracle).

Pattern and order must be identical; block values agree to 1e-13 Ha (device
and host libm may differ in the last ulp of acos / cos / exp)."""
import numpy as np
import pytest

import largetb_amd as L

pytestmark = pytest.mark.gpu
BOHR = 1.88973


def quartz_cell(l, m, n, lattice=L.LATTICE_CONFIG):
    a, c = (9.725, 10.7) if lattice == L.LATTICE_REFERENCE else (4.913 * BOHR, 5.405 * BOHR)
    R = np.array([[a, 0, 0], [-a * np.sin(np.pi / 6), a * np.cos(np.pi / 6), 0], [0, 0, c]])
    return np.array([l * R[0], m * R[1], n * R[2]])


def same_system(ctx, s):
    rp, col, val = s.hamiltonian()
    drp, dcol, dval = ctx.hamiltonian()
    assert np.array_equal(rp, drp) and np.array_equal(col, dcol)
    assert np.abs(val - dval).max() < 1e-13
    grp, gcol, gval = s.gamma()
    dgrp, dgcol, dgval = ctx.gamma()
    assert np.array_equal(grp, dgrp) and np.array_equal(gcol, dgcol)
    assert np.abs(gval - dgval).max() < 1e-13
    assert ctx.nelectrons == s.nelectrons


@pytest.mark.parametrize("cells,sigma,lattice", [((2, 2, 2), 0.05, L.LATTICE_CONFIG),
                                                  ((8, 8, 8), 0.05, L.LATTICE_CONFIG),
                                                  ((5, 3, 4), 0.15, L.LATTICE_CONFIG),
                                                  ((4, 4, 4), 0.0, L.LATTICE_REFERENCE)])
def test_device_assembly_equals_host(ctx, cells, sigma, lattice):
    s = L.System.quartz(*cells, lattice=lattice, sigma_ang=sigma, seed=3)
    Z, xyz = s.coords()
    mask = 3 if lattice == L.LATTICE_REFERENCE else 7
    ctx.load_coords(Z, xyz, quartz_cell(*cells, lattice), periodic_mask=mask)
    same_system(ctx, s)


def test_device_assembly_cluster_without_cell(ctx):
    """No periodicity: Cartesian bins over the bounding box."""
    s0 = L.System.quartz(3, 3, 3, sigma_ang=0.05, seed=2)
    Z, xyz = s0.coords()
    s = L.System.from_coords(Z, xyz, None, periodic_mask=0)
    ctx.load_coords(Z, xyz, None, periodic_mask=0)
    same_system(ctx, s)


def test_device_assembly_large_cell(ctx):
    """207,360 atoms (24 x 24 x 40 cells)."""
    s = L.System.quartz(24, 24, 40, sigma_ang=0.08, seed=5)
    Z, xyz = s.coords()
    sec = ctx.load_coords(Z, xyz, quartz_cell(24, 24, 40))
    assert sec > 0
    same_system(ctx, s)


def test_scc_on_device_assembled_system(ctx):
    s = L.System.quartz(3, 3, 2, sigma_ang=0.1, seed=4)
    ctx.load(s)
    host = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    Z, xyz = s.coords()
    ctx.load_coords(Z, xyz, quartz_cell(3, 3, 2))
    dev = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    assert dev["iters"] == host["iters"]
    assert np.abs(dev["q"] - host["q"]).max() < 1e-10


def test_device_assembly_errors(ctx):
    s = L.System.quartz(2, 2, 2, sigma_ang=0.0, seed=1)
    Z, xyz = s.coords()
    # a periodic row narrower than 8 Bohr (one image per pair not guaranteed)
    with pytest.raises(L.LtbError) as e:
        ctx.load_coords(Z, xyz, np.diag([7.5, 20.0, 20.0]))
    assert e.value.status == L.LTB_ERR_ARG
    # two atoms on top of each other
    bad = xyz.copy()
    bad[1] = bad[0] + 0.1
    with pytest.raises(L.LtbError) as e:
        ctx.load_coords(Z, bad, quartz_cell(2, 2, 2))
    assert e.value.status == L.LTB_ERR_GEOMETRY
    # the context stays usable
    ctx.load_coords(Z, xyz, quartz_cell(2, 2, 2))
    same_system(ctx, s)
