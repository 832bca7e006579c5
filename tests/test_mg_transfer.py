"""The multigrid transfers (CPU): the prolongation built from the 1-D
weights one unknown at a time equals the Kronecker construction, and the
interpolation properties the weights encode hold (constants kept away from
the no-slip walls, zero inflow at the walls)."""
import numpy as np
import pytest

import mg_ref


@pytest.mark.parametrize("nxc,nyc", [(4, 3), (8, 4), (5, 6)])
def test_prolongation_matches_kron_construction(nxc, nyc):
    a = mg_ref.prolongation(nxc, nyc).toarray()
    b = mg_ref.prolongation_kron(nxc, nyc).toarray()
    np.testing.assert_array_equal(a, b)


def test_prolongation_interpolates():
    nxc, nyc = 8, 6
    P = mg_ref.prolongation(nxc, nyc)
    mnc, mnf = (nxc + 1) * nyc, (2 * nxc + 1) * 2 * nyc
    # v = 1 everywhere: the fine v is 1 (cells mirror at the side walls; every fine inner face is
    # interpolated from inner coarse faces, or lies between an inner face and a wall)
    xc = np.zeros(P.shape[1])
    xc[mnc:] = 1.0
    xf = P @ xc
    vf = xf[mnf:].reshape(2 * nyc - 1, 2 * nxc)
    np.testing.assert_allclose(vf[1:-1], 1.0)  # away from the wall rows
    np.testing.assert_allclose(vf[[0, -1]], 0.5)  # halfway between a wall (v = 0) and a face
    # u linear in x is reproduced exactly along x (faces interpolate linearly)
    xc = np.zeros(P.shape[1])
    I = np.tile(np.arange(nxc + 1), nyc)
    xc[:mnc] = I
    uf = (P @ xc)[:mnf].reshape(2 * nyc, 2 * nxc + 1)
    np.testing.assert_allclose(uf[2:-2], np.tile(np.arange(2 * nxc + 1) / 2.0, (2 * nyc - 4, 1)))
