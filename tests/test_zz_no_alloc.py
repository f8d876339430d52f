"""spx_gradre_noalloc (work arrays in the library's static device scratch)
against spx_gradre (work arrays the library allocates), through the C ABI:
bit-identical results on two calls in a row, the second on a smaller grid,
in both precisions; grids beyond the scratch are refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(plan, seed):
    import torch
    rng = np.random.default_rng(seed)
    shape = plan.spec_shape
    return [torch.as_tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).cuda()
            for _ in range(3)]


def _call(plan, fn, v):
    from specter_amd import lib as L
    outs = [plan.empty_nonlinear() for _ in range(3)]
    args = [ctypes.c_void_p(t.data_ptr()) for t in (*v, *outs)]
    L.check(fn(plan.handle, *args))
    plan.sync()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_gradre_noalloc_matches_gradre(precision):
    from specter_amd import FCPlan
    from specter_amd.lib import load_library
    lib = load_library()
    for shape, seed in (((32, 24, 89, 25, 5), 1), ((16, 16, 41, 25, 5), 2)):
        p_ref = FCPlan(*shape, precision=precision)
        p_na = FCPlan(*shape, precision=precision)
        need, limit = ctypes.c_size_t(), ctypes.c_size_t()
        assert lib.spx_gradre_scratch_bytes(p_na.handle, ctypes.byref(need),
                                            ctypes.byref(limit)) == 0
        assert 0 < need.value <= limit.value
        v = _fields(p_ref, seed)
        ref = _call(p_ref, lib.spx_gradre, v)
        got = _call(p_na, lib.spx_gradre_noalloc, v)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), shape
        # the no-alloc plan never allocated its own work arrays
        assert p_na.info().work_bytes > 0
        p_ref.close()
        p_na.close()


def test_gradre_noalloc_refuses_large_grids():
    import torch
    from specter_amd import FCPlan
    from specter_amd.lib import SpxError, load_library
    lib = load_library()
    p = FCPlan(256, 256, 281, 25, 5)
    need, limit = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.spx_gradre_scratch_bytes(p.handle, ctypes.byref(need), ctypes.byref(limit)) == 0
    assert need.value > limit.value
    dummy = torch.zeros(1, dtype=torch.complex128, device="cuda")
    ptr = ctypes.c_void_p(dummy.data_ptr())
    rc = lib.spx_gradre_noalloc(p.handle, ptr, ptr, ptr, ptr, ptr, ptr)
    assert rc != 0
    buf = ctypes.create_string_buffer(256)
    lib.spx_last_error(buf, 256)
    assert b"static scratch" in buf.value
    p.close()
