"""The allocation-free batched dot (mq_vec_multi_dot_noalloc), called through the C-ABI, equals the
allocating mq_vec_multi_dot bit for bit (same kernels, same fixed reduction order), on two calls in
a row (the second with fewer and shorter vectors, over the scratch the first left behind), and
rejects k beyond its static scratch before touching the device."""
import ctypes as C

import numpy as np
import pytest
import torch

from mqdmrg import engine as E
from mqdmrg.capi import lib

pytestmark = pytest.mark.gpu

MAXK = 64   # MQ_NOALLOC_MAXK of include/mq_sigma.h


def _noalloc(eng, V, k, ldv, x, res):
    n = x.numel()
    return lib().mq_vec_multi_dot_noalloc(eng.ctx, n, int(k), E.dev_ptr(eng, V, (k - 1) * ldv + n), int(ldv),
                                          E.dev_ptr(eng, x, n), E.dev_ptr(eng, res, k))


def test_noalloc_multi_dot_equals_allocating_form():
    eng = E.Engine(0)
    try:
        g = torch.Generator(device="cpu").manual_seed(11)
        for k, n, ldv in ((MAXK, 300_001, 300_032), (5, 1_000, 1_024)):
            print("case k=%d n=%d ldv=%d" % (k, n, ldv), flush=True)
            V = torch.randn(k * ldv, dtype=torch.float64, generator=g).to(eng.torch_device)
            x = torch.randn(n, dtype=torch.float64, generator=g).to(eng.torch_device)
            res = torch.full((k,), float("nan"), dtype=torch.float64, device=eng.torch_device)
            assert _noalloc(eng, V, k, ldv, x, res) == 0, lib().mq_last_error()
            eng.sync()
            got = res.cpu().numpy()
            ref = eng.multi_dot(V, k, ldv, x)
            np.testing.assert_array_equal(got, ref)
            Vh = V.cpu().numpy()
            xh = x.cpu().numpy()
            exact = np.array([Vh[j * ldv:j * ldv + n] @ xh for j in range(k)])
            assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max()
        # k beyond the static scratch is an error, not an allocation
        V = torch.zeros((MAXK + 1) * 8, dtype=torch.float64, device=eng.torch_device)
        x = torch.zeros(8, dtype=torch.float64, device=eng.torch_device)
        res = torch.zeros(MAXK + 1, dtype=torch.float64, device=eng.torch_device)
        assert _noalloc(eng, V, MAXK + 1, 8, x, res) != 0
        assert b"MQ_NOALLOC_MAXK" in lib().mq_last_error()
    finally:
        eng.close()
