"""The allocation-free direct sum (ffmm_direct_noalloc) through the C ABI, against
ffmm_direct on the same inputs: two calls in a row (the second smaller, so that the
fixed device array still holds the first call's partial sums), bitwise equal outputs,
both within 1e-12 of the oracle's FP64 direct sum; targets beyond the array's limit are
refused before anything is launched."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20        # targets the 32 MiB array holds partial sums for


def _call(name, ctx, s, t, x, sk, out, code):
    from ffmm import _lib
    p = lambda a: None if a is None else ctypes.c_void_p(a.data_ptr())   # noqa: E731
    return getattr(_lib.lib(), name)(ctx, p(s), p(t), s.shape[0], p(x), x.shape[0], p(sk), p(out), code)


@pytest.mark.parametrize("output,ncomp", [(0, 1), (2, 4)])
def test_noalloc_direct_equals_direct(output, ncomp):
    import torch
    from ffmm import _lib
    from ffmm.fmm import context
    ctx = context(0)
    rng = np.random.default_rng(70 + output)
    dev = "cuda:0"
    for ns, nt, with_skip in ((30_000, 5_000, True), (4_000, 700, False)):
        print(f"[no_alloc] output={output} ns={ns} nt={nt} skip={with_skip}", flush=True)
        src = rng.random((ns, 3))
        tau = rng.standard_normal(ns)
        if with_skip:
            rows = np.sort(rng.choice(ns, nt, replace=False))
            tgt = src[rows]
            skip = rows.astype(np.int64)
        else:
            tgt = rng.random((nt, 3))
            skip = None
        s = torch.from_numpy(src).to(dev)
        t = torch.from_numpy(tau).to(dev)
        x = torch.from_numpy(np.ascontiguousarray(tgt)).to(dev)
        sk = None if skip is None else torch.from_numpy(skip).to(dev)
        a = torch.full((nt, ncomp), np.nan, dtype=torch.float64, device=dev)
        b = torch.full((nt, ncomp), np.nan, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        assert _call("ffmm_direct", ctx, s, t, x, sk, a, output) == _lib.FFMM_OK
        assert _call("ffmm_direct_noalloc", ctx, s, t, x, sk, b, output) == _lib.FFMM_OK, \
            _lib.lib().ffmm_last_error().decode()
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.array_equal(a, b)
        pot = O.direct(src, tau, tgt) if skip is None else O.direct_self(src, tau, rows=skip)
        assert np.max(np.abs(b[:, 0] - pot)) <= 1e-12 * np.max(np.abs(pot))
        if ncomp == 4:
            g = O.direct(src, tau, tgt, grad=True) if skip is None else O.direct_self(src, tau, grad=True, rows=skip)
            assert np.max(np.abs(b[:, 1:] - g)) <= 1e-12 * np.max(np.abs(g))


def test_noalloc_direct_refuses_targets_beyond_its_array():
    import torch
    from ffmm import _lib
    from ffmm.fmm import context
    ctx = context(0)
    dev = "cuda:0"
    s = torch.zeros((16, 3), dtype=torch.float64, device=dev)
    t = torch.ones(16, dtype=torch.float64, device=dev)
    x = torch.zeros((LIMIT + 1, 3), dtype=torch.float64, device=dev)
    out = torch.zeros(LIMIT + 1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print("[no_alloc] refusal above the limit", flush=True)
    assert _call("ffmm_direct_noalloc", ctx, s, t, x, None, out, 0) == _lib.FFMM_ERR_ARGUMENT
    assert "at most 1048576 targets" in _lib.lib().ffmm_last_error().decode()
    assert float(out.abs().max()) == 0.0          # nothing was launched
