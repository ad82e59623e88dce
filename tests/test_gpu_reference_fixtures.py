"""The engine's two-site DMRG and sigma adjoint against the reference's own regression observables
(tests/golden/reference_observables.json, exact diagonalisation results stored with the reference's
tests): the fermion Hubbard ladder of test_coded_fermiladder_ts.py (energy, local densities, the
one-body density matrix, the latter two as dE/dh from the adjoint) and the 5 x 5 free fermions of
test_ff_ts.py (energy).  Both run the TwoU1 ab initio MPO of the lattice's integrals."""
import numpy as np
import pytest

from lattices import reference_observables, ladder_hopping, square_hopping, hubbard_integrals
from mqdmrg import engine as E, qc
from mqdmrg.dmrg import TwoSiteDMRG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def test_fermiladder_energy_densities_and_onebody_matrix(eng):
    """Energy to 1e-8 of the exact 55.08693016528693; the adjoint's dE/dh_pq, the spin-summed
    one-body density matrix <c+_p c_q>, equals the fixture's up + down matrices to 1e-6 (its
    diagonal: the local densities)."""
    ref = reference_observables()["fermiladder"]
    h, eri = hubbard_integrals(ladder_hopping(6), 8.0)
    # every one-body entry is a term (zeros included), so dE/dh covers the whole density matrix
    basis, mpo, _, src = qc.qc_mpo(h, eri, sources=True, h_tol=-1.0)
    d = TwoSiteDMRG(eng, "qc", 12, m=512, cutoff=1e-14, init_m=16, seed=3, jd_tol=1e-12, jd_maxiter=20,
                    mpo=mpo, basis=basis, target=(10, 10))
    e = d.run(4)
    print("E %.12f ref %.12f" % (e, ref["energy"]), flush=True)
    assert abs(e - ref["energy"]) <= 1e-8
    grads, energies = d.mpo_gradients()
    assert max(abs(x - e) for x in energies) < 1e-8
    dh, _, _ = qc.mpo_gradient_to_integrals(mpo, src, grads, 12)
    gamma = np.array(ref["onebody_up"]) + np.array(ref["onebody_down"])
    dens = np.array(ref["local_density_up"]) + np.array(ref["local_density_down"])
    err = np.abs(dh - gamma).max()
    print("max |dE/dh - gamma| = %.3e, max density error %.3e" % (err, np.abs(np.diag(dh) - dens).max()), flush=True)
    assert err <= 1e-6
    assert np.abs(np.diag(dh) - dens).max() <= 1e-6


def test_free_fermions_energy(eng):
    """Spinless fermions (the up species of the ab initio basis, no down electrons) on the open 5 x 5
    lattice, N=10: the reference's test runs m=200 against exact diagonalisation at 1e-4; here m=256
    against the exact -17.856406460551096 at 1e-5."""
    ref = reference_observables()["free_fermions"]
    h = square_hopping(5, 5)
    eri = np.zeros((25,) * 4)
    basis, mpo, _ = qc.qc_mpo(h, eri)
    d = TwoSiteDMRG(eng, "qc", 25, m=256, cutoff=1e-12, init_m=16, seed=5, jd_tol=1e-10, jd_maxiter=12,
                    mpo=mpo, basis=basis, target=(10, 0))
    e = d.run(4)
    print("E %.12f ref %.12f diff %.3e" % (e, ref["energy"], e - ref["energy"]), flush=True)
    assert abs(e - ref["energy"]) <= 1e-5
