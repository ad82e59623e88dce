"""CPU checks of the reference's exact-diagonalisation fixtures (tests/golden/reference_observables.json,
written from the reference's regression HDF5 files) against independent exact solutions of the
lattice models as this repository writes them (site numbering, hopping signs, integrals), so that
the GPU DMRG tests compare against fixtures whose meaning is pinned."""
import numpy as np

from lattices import (reference_observables, ladder_hopping, square_hopping, hubbard_integrals,
                      exact_hubbard)
from mqdmrg import qc


def test_fermiladder_fixture_equals_exact_solution():
    """Hubbard ladder L=6, t=t'=1, U=8, 10+10 electrons: energy, local densities and both spins'
    one-body density matrices of the fixture equal a direct diagonalisation of the ladder
    (site = 2 * rung + leg)."""
    ref = reference_observables()["fermiladder"]
    e, gu, gd = exact_hubbard(ladder_hopping(6), 8.0, 10, 10)
    assert abs(e - ref["energy"]) <= 1e-10
    assert abs(e - ref["dmrg_test_energy"]) <= 1e-10
    np.testing.assert_allclose(np.diag(gu), ref["local_density_up"], atol=1e-9)
    np.testing.assert_allclose(np.diag(gd), ref["local_density_down"], atol=1e-9)
    np.testing.assert_allclose(gu, np.array(ref["onebody_up"]), atol=1e-9)
    np.testing.assert_allclose(gd, np.array(ref["onebody_down"]), atol=1e-9)


def test_free_fermion_fixture_equals_closed_form():
    """Spinless free fermions on the open 5 x 5 lattice, N=10: the fixture's energy is the sum of
    the 10 lowest single-particle energies -2 cos(pi a / 6) - 2 cos(pi b / 6)."""
    ref = reference_observables()["free_fermions"]
    k = np.pi * np.arange(1, 6) / 6.0
    eps = np.sort((-2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :]).ravel())
    assert abs(eps[:10].sum() - ref["energy"]) <= 1e-12
    lat = np.sort(np.linalg.eigvalsh(square_hopping(5, 5)))
    assert abs(lat[:10].sum() - ref["energy"]) <= 1e-12
    assert lat[10] - lat[9] > 0.5   # closed shell: a unique ground state


def test_hubbard_integrals_reproduce_the_model():
    """The ab initio form of the Hubbard model ((pp|pp) = U) has the Hubbard spectrum: the FCI of
    those integrals equals the direct diagonalisation on a small ladder (2 rungs, 2+1 electrons)."""
    hop = ladder_hopping(2)
    h, eri = hubbard_integrals(hop, 8.0)
    e_fci = qc.fci_energy(h, eri, 2, 1)
    e_ex, _, _ = exact_hubbard(hop, 8.0, 2, 1)
    assert abs(e_fci - e_ex) <= 1e-12
