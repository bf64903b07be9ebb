"""Geometry helpers used to build the pointer arrays of the benchmarks/tests."""
import numpy as np

from qla_amd import lattice


def test_neighbours_inverse_and_parity():
    dims = (4, 6, 2, 8)
    n = lattice.volume(dims)
    fw, bw = lattice.neighbours(dims)
    for mu in range(4):
        np.testing.assert_array_equal(fw[mu][bw[mu]], np.arange(n))
        np.testing.assert_array_equal(bw[mu][fw[mu]], np.arange(n))
    even, odd = lattice.parity_sites(dims, 0), lattice.parity_sites(dims, 1)
    assert len(even) == len(odd) == n // 2
    par = lattice.coords(dims).sum(axis=1) % 2
    for mu in range(4):
        assert (par[fw[mu][even]] == 1).all()


def test_lexicographic_order():
    dims = (3, 4, 5, 6)
    c = lattice.coords(dims)
    np.testing.assert_array_equal(lattice.site_index(c, dims), np.arange(lattice.volume(dims)))
    assert tuple(c[1]) == (1, 0, 0, 0) and tuple(c[3]) == (0, 1, 0, 0)


def test_su3_is_special_unitary():
    u = lattice.random_su3(200, 1)
    z = u[..., 0] + 1j * u[..., 1]
    eye = np.einsum("nij,nkj->nik", z, z.conj())
    assert np.abs(eye - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(z) - 1).max() < 1e-13


def test_seeded_reproducible():
    a = lattice.random_field("D", 100, 5)
    b = lattice.random_field("D", 100, 5)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, lattice.random_field("D", 100, 6))
