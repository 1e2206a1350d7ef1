"""m3p_mhd_advance_ct_noalloc (the entry point that allocates nothing, its
MHD workspace a static device array) against m3p_mhd_advance_ct with a
caller workspace: the same kernels, so every output and the time step must
be equal bit for bit, on calls in a row, each on a smaller cube
(the static array keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _state(N, seed):
    from mara3p_amd.problems import orszag_tang_3d
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=seed, perturbation=0.05)
    return [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in (uc, b1, b2, b3)]


def _advance(lib, src, base, N, precision, noalloc):
    from mara3p_amd._lib import check
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = [torch.full_like(t, np.nan) for t in src]
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    b = base if base is not None else [None] * 4
    args = [_p(src[0]), _p(src[1]), _p(src[2]), _p(src[3]), _p(b[0]), _p(b[1]), _p(b[2]), _p(b[3]), 0.5,
            N, 1.0 / N, 0.33, 1.5, 5.0 / 3.0, _p(out[0]), _p(out[1]), _p(out[2]), _p(out[3]), _p(dt)]
    if noalloc:
        status = torch.full((4,), 7, dtype=torch.int32, device="cuda")
        check(lib.m3p_mhd_advance_ct_noalloc(*args, _p(status), precision, stream))
        torch.cuda.synchronize()
        assert status.cpu().tolist() == [0, 0, 0, 0]
    else:
        work = torch.empty(lib.m3p_mhd_workspace_bytes(N, precision), dtype=torch.uint8, device="cuda")
        check(lib.m3p_mhd_advance_ct(*args, _p(work), work.numel(), precision, stream))
        torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out], float(dt.item())


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_mhd_equals_workspace_entry(precision):
    from mara3p_amd._lib import lib
    L = lib()
    # later calls on smaller cubes; 20 takes the per-cell passes, 40 and 24
    # the column tiles
    for N, seed in ((40, 1), (24, 2), (20, 3)):
        src = _state(N, seed)
        base = _state(N, seed + 10)
        for b in (None, base):               # first stage, then an RK combine
            ref, dt_ref = _advance(L, src, b, N, precision, noalloc=False)
            got, dt_got = _advance(L, src, b, N, precision, noalloc=True)
            assert dt_got == dt_ref
            for x, y in zip(got, ref):
                assert np.array_equal(x, y)


def test_noalloc_mhd_rejects_too_large_a_cube():
    """A cube whose workspace exceeds the static array is refused before any
    launch, with the largest N that fits named in the message."""
    from mara3p_amd._lib import lib
    L = lib()
    N = 96
    bufs = [torch.zeros(8, dtype=torch.float64, device="cuda") for _ in range(8)]
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    ins, outs = [_p(t) for t in bufs[:4]], [_p(t) for t in bufs[4:]]
    st = L.m3p_mhd_advance_ct_noalloc(*ins, None, None, None, None, 0.5, N, 1.0 / N, 0.33, 1.5, 5.0 / 3.0,
                                      *outs, None, _p(status), 0,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st != 0
    assert b"static scratch" in L.m3p_last_error()
