"""qew_eval_noalloc (no-allocation variant of qew_eval) through the C ABI:
its gradient and energy equal qew_eval's bit for bit, on two calls in a row
(the second on a shorter line, so stale partials of the first call would
show), and lines beyond its fixed scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, host, torch_cuda

pytestmark = pytest.mark.gpu


def test_eval_noalloc_matches_eval():
    torch = torch_cuda()
    import qew_amd as q
    from qew_amd import _lib
    lib = _lib.load()
    ctx = q.context(0)
    rng = np.random.default_rng(11)
    ny = 512
    for n, model, seed in ((70000, 1, 3), (3000, 0, 4)):
        field, coef = O.filtered_noise(n, ny, float(n), float(ny), 1.0, 4.0, 4.0, seed)
        noise = q.FilteredNoise.from_coefficients(coef, 1.0, 1.0)
        p = q.Params(float(n), 1.5, 0.1, q.Model(model))
        pc = p.c()
        h = dev(ny / 2 + np.cumsum(rng.standard_normal(n)))
        e_ref, g_ref = q.evaluate(h, p, noise)
        grad = torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0")
        energy = torch.full((1,), np.nan, dtype=torch.float64, device="cuda:0")
        _lib.check(lib.qew_eval_noalloc(ctx.handle, noise.handle, C.byref(pc), n,
                                        C.c_void_p(h.data_ptr()), C.c_void_p(grad.data_ptr()),
                                        C.c_void_p(energy.data_ptr())))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host(grad), host(g_ref))
        assert float(energy[0]) == e_ref, (n, float(energy[0]), e_ref)
        # gradient only
        grad2 = torch.zeros_like(grad)
        _lib.check(lib.qew_eval_noalloc(ctx.handle, noise.handle, C.byref(pc), n,
                                        C.c_void_p(h.data_ptr()), C.c_void_p(grad2.data_ptr()),
                                        None))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host(grad2), host(g_ref))
        # and the oracle agrees (FP64 gradient bit for bit, energy to 1e-12)
        po = O.params(float(n), 1.5, 0.1, model)
        on = O.OracleNoise.table(coef, float(n), float(ny))
        np.testing.assert_array_equal(host(grad), O.gradient(host(h), po, on))
        assert abs(float(energy[0]) - O.objective(host(h), po, on)) <= 1e-12 * abs(e_ref)


def test_eval_noalloc_rejects_lines_beyond_its_scratch():
    torch_cuda()
    import qew_amd as q
    from qew_amd import _lib
    lib = _lib.load()
    ctx = q.context(0)
    noise = q.ZeroNoise()
    pc = q.Params(1.0, 1.0, 0.0).c()
    n = 16384 * 255 + 1
    st = lib.qew_eval_noalloc(ctx.handle, noise.handle, C.byref(pc), n, C.c_void_p(16),
                              C.c_void_p(16), None)
    assert st == _lib.QEW_ERR_INVALID_ARGUMENT
    assert b"maximum is 16384 tiles" in lib.qew_last_error()
