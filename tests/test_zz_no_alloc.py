"""ngw_qp_newton_dev_noalloc -- the entry point that allocates nothing (its
spline work arrays are a fixed per-GPU device array) -- against the entry it
varies, ngw_qp_newton, through the C ABI: two calls in a row, the second on a
smaller input (fewer states, fewer grid points), each equal to ngw_qp_newton
on the same data (same kernel, same inputs: bitwise); a size above the
stated limit is NGW_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tables(rng, nqp, nen):
    x = np.linspace(-1, 1, nen)
    sig = np.empty((nen, nqp), np.complex128)
    for i in range(nqp):
        c = rng.normal(size=4)
        sig[:, i] = (c[0] + c[1] * x + c[2] * np.sin(3 * x) + 0.3 * c[3] * x ** 3) + 1j * (0.1 * np.cos(x))
    edft = rng.uniform(-10, 0, nqp)
    sigx = rng.uniform(-5, 0, nqp)
    vxc = rng.uniform(-3, 0, nqp)
    return sig * 13.6, edft, sigx, vxc


def _noalloc(ctx, gpu, sig, edft, sigx, vxc, deltae):
    import torch

    from nanogw_amd import _lib

    nen, nqp = sig.shape
    head = np.stack([edft, edft, sigx, vxc], axis=1).reshape(-1)
    packed = np.concatenate([head, np.asfortranarray(sig).reshape(-1, order="F").view(np.float64)])
    d_in = torch.tensor(packed, dtype=torch.float64, device=gpu)
    d_out = torch.full((6 * nqp,), np.nan, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    lib = _lib.load()
    st = lib.ngw_qp_newton_dev_noalloc(ctx.handle, nqp, nen, C.c_double(deltae), C.c_void_p(d_in.data_ptr()),
                                       C.c_void_p(d_out.data_ptr()))
    _lib.check(st, "ngw_qp_newton_dev_noalloc")
    _lib.check(lib.ngw_synchronize(ctx.handle), "ngw_synchronize")
    return d_out.cpu().numpy().reshape(nqp, 6)


def test_qp_newton_noalloc_equals_qp_newton(gpu, ngw_ctx):
    rng = np.random.default_rng(2027)
    deltae = 27.2
    for nqp, nen in [(40, 200), (7, 61)]:
        sig, edft, sigx, vxc = _tables(rng, nqp, nen)
        ref = ngw_ctx.qp_newton(sig, deltae, edft, edft, sigx, vxc)
        out = _noalloc(ngw_ctx, gpu, sig, edft, sigx, vxc, deltae)
        print(f"nqp={nqp} nen={nen}: max |dE| = {np.abs(out[:, 0] - ref['eqp']).max():.3e}")
        assert np.array_equal(out[:, 0], ref["eqp"])
        assert np.array_equal(out[:, 1], ref["zfac"])
        assert np.array_equal(out[:, 2], ref["sigma0"].real) and np.array_equal(out[:, 3], ref["sigma0"].imag)
        assert np.array_equal(out[:, 4], ref["imsig"])
        assert np.array_equal(out[:, 5].astype(np.int32), ref["ierr"])


def test_qp_newton_noalloc_rejects_beyond_limit(gpu, ngw_ctx):
    """257 x 256 points is one grid row past the limit: refused on the host
    before any work is queued."""
    import torch

    from nanogw_amd import _lib

    lib = _lib.load()
    d = torch.zeros(8, dtype=torch.float64, device=gpu)
    st = lib.ngw_qp_newton_dev_noalloc(ngw_ctx.handle, 257, 256, C.c_double(27.2), C.c_void_p(d.data_ptr()),
                                       C.c_void_p(d.data_ptr()))
    assert st == -1
    assert "NGW_QP_NOALLOC_MAX_POINTS" in lib.ngw_last_error().decode()
