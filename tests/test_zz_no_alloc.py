"""pw_nonlocal_multx_noalloc (the allocation-free NonlocalMultX, its
scratch in a fixed per-device array) against pw_nonlocal_multx through the
C-ABI: two calls in a row, the second with fewer bands, FP64 and FP32; a
shape whose split-K partials do not all fit the scratch (fewer K splits:
FP64 rounding level); a shape too large for the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from pwdft_develop_amd import _lib
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


def _noalloc(ctx, x, out):
    from pwdft_develop_amd.ops import _dtype_code
    return _lib.lib().pw_nonlocal_multx_noalloc(ctx.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                                C.c_void_p(out.data_ptr()))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_matches_nonlocal_multx(si8, dtype):
    import torch
    from pwdft_develop_amd.ops import Context
    print(f"case: si8_32 nonlocal_multx_noalloc {dtype}", flush=True)
    tdt = torch.complex128 if dtype == "f64" else torch.complex64
    ctx = Context(si8)
    for nb in (24, 7):
        X = ctx.to_device(si8.psi[:nb], tdt)
        ref = ctx.nonlocal_multx(X)
        out = torch.full_like(X, float("nan"))
        _lib.check(_noalloc(ctx, X, out))
        torch.cuda.synchronize()
        # the usual split count fits the scratch here: the same products in
        # the same order
        assert torch.equal(out, ref), (nb, float((out - ref).abs().max()))
    ctx.close()


def test_noalloc_fewer_splits_than_usual():
    """si64_60, 256 bands, FP64: 64 usual K splits of 512 KiB partials do
    not fit next to W in 32 MiB, so the call takes fewer: FP64 agreement."""
    import torch
    from pwdft_develop_amd.ops import Context
    print("case: si64_60 nonlocal_multx_noalloc fp64 nb=256 then 100", flush=True)
    prob = W.si64_60(nb=256).build()
    ctx = Context(prob)
    for nb in (256, 100):
        X = ctx.to_device(prob.psi[:nb])
        ref = ctx.nonlocal_multx(X)
        out = torch.empty_like(X)
        _lib.check(_noalloc(ctx, X, out))
        torch.cuda.synchronize()
        err = float((out - ref).abs().max() / ref.abs().max())
        assert err < 1e-13, (nb, err)
    ctx.close()


def test_noalloc_refuses_oversized_projections(si8):
    """4096 projectors x 1024 bands in FP64 (32 MiB of projections alone)
    exceed the scratch: PW_ERR_UNSUPPORTED with a message, nothing run."""
    import torch
    from pwdft_develop_amd.ops import Context
    print("case: si8_32 nonlocal_multx_noalloc oversized", flush=True)
    ctx = Context(si8, max_bands=1024)
    nbeta = 4096
    rng = np.random.default_rng(3)
    beta = rng.standard_normal(2 * si8.npw * nbeta)
    w = rng.uniform(0.5, 2.0, nbeta)
    _lib.check(_lib.lib().pw_set_nonlocal(ctx.h, nbeta, beta, w))
    X = torch.zeros((1024, si8.npw), dtype=torch.complex128, device="cuda")
    out = torch.empty_like(X)
    rc = _noalloc(ctx, X, out)
    assert rc == -7 and b"scratch" in _lib.lib().pw_last_error()
    ctx.close()
