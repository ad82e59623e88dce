"""Boundary residency on the GPU (mq_boundary_evict / _prefetch / _residency): evicted, spilled
and prefetched boundaries give bitwise the same sigma and data; a boundary a live plan reads cannot
be evicted; two-site DMRG with the boundary cache gives the same energies as without it while
keeping only a few boundaries on the device."""
import numpy as np
import pytest
import torch

from chain import SiteProblem
from mqdmrg import capi, engine as E
from mqdmrg.dmrg import TwoSiteDMRG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = E.Engine(0)
    yield e
    e.close()


def test_evict_spill_prefetch_round_trip(eng, tmp_path):
    p = SiteProblem("hubbard", L=8, site=3, Mmax=8, seed=5)
    L = E.DeviceBoundary.upload(eng, p.Lb.index, p.Lb.cohorts)
    R = E.DeviceBoundary.upload(eng, p.Rb.index, p.Rb.cohorts)
    psi = torch.from_numpy(p.psi).to(eng.torch_device)

    def sigma():
        plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, L, R, algorithm=E.SIGMA_GROUPED)
        out = torch.zeros_like(psi)
        plan.apply(psi, out)
        eng.sync()
        plan.close()
        return out.clone()

    ref = sigma()
    ref_data = L.download()
    where, dev_bytes, _ = L.residency()
    assert where == E.DeviceBoundary.DEVICE and dev_bytes > 0
    # a live plan pins its boundaries
    plan = E.SigmaPlan(eng, p.site_desc, p.mpo_desc, L, R, algorithm=E.SIGMA_GROUPED)
    with pytest.raises(capi.EngineError):
        L.evict()
    plan.close()
    # to host: the device copy is freed; a plan brings it back by itself
    L.evict()
    R.evict()
    where, dev_bytes, host_bytes = L.residency()
    assert where == E.DeviceBoundary.HOST and dev_bytes == 0 and host_bytes > 0
    assert torch.equal(sigma(), ref)
    # to disk: the host copy is freed too; an explicit prefetch, then the same data and sigma
    path = tmp_path / "left.bin"
    L.evict(path)
    where, dev_bytes, host_bytes = L.residency()
    assert where == E.DeviceBoundary.DISK and dev_bytes == 0 and host_bytes == 0 and path.exists()
    L.prefetch()
    assert L.residency()[0] == E.DeviceBoundary.DEVICE
    for a, b in zip(L.download(), ref_data):
        assert np.array_equal(a, b)
    assert torch.equal(sigma(), ref)
    # evicting again keeps the host copy already made (no new copy) and is repeatable
    L.evict()
    L.evict()
    assert torch.equal(sigma(), ref)
    L.close()
    assert not path.exists()


@pytest.mark.parametrize("spill", [False, True])
def test_dmrg_with_boundary_cache_matches(eng, tmp_path, spill):
    """Hubbard L=10, m=64, 3 sweeps: per-sweep energies equal those of the run without the cache to
    1e-12; after the run at most 5 boundaries are on the device (without the cache: all 2L)."""
    print("case spill", spill, flush=True)
    kw = dict(m=64, cutoff=1e-12, init_m=8, seed=4, jd_tol=1e-10, U=4.0)
    a = TwoSiteDMRG(eng, "hubbard", 10, **kw)
    a.run(3)
    b = TwoSiteDMRG(eng, "hubbard", 10, boundary_cache=True, spill_dir=str(tmp_path) if spill else None, **kw)
    b.run(3)
    print("energies", a.energies, b.energies, "resident", a.resident_boundaries(), b.resident_boundaries(), flush=True)
    assert max(abs(x - y) for x, y in zip(a.energies, b.energies)) <= 1e-12
    assert b.resident_boundaries() <= 5 < a.resident_boundaries()
