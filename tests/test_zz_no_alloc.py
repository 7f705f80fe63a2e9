"""lbm_reduce_f64_static, the reduction that allocates nothing (its partial
sums live in a static device array of the library), against lbm_reduce_f64
with a caller workspace, through the C-ABI: bit for bit (both run the same
fixed-order kernels), on two calls in a row with the second on a smaller
grid, with and without an exclusion mask."""
import ctypes

import pytest
import torch

import lettuce_amd as lt
from lettuce_amd import _native as N
from lettuce_amd._ops import STORAGE_OF, layout

pytestmark = pytest.mark.gpu


def _both(flow, mask=None):
    f = flow._state
    S = flow.stencil
    n, s = layout(f, S.q, S.d)
    n3, s4 = (ctypes.c_int64 * 3)(*n), (ctypes.c_int64 * 4)(*s)
    stream = N.stream_handle(f.device)
    mask_ptr = None
    if mask is not None:
        mask = mask.to(device=f.device, dtype=torch.uint8).contiguous()
        mask_ptr = mask.data_ptr()
    ws = torch.empty(N.lib().lbm_reduce_workspace_bytes(n3) // 8, dtype=torch.float64,
                     device=f.device)
    out_ws = torch.full((5,), float("nan"), dtype=torch.float64, device=f.device)
    out_st = torch.full((5,), float("nan"), dtype=torch.float64, device=f.device)
    N.check(N.lib().lbm_reduce_f64(S.native_id, STORAGE_OF[f.dtype], n3, s4, f.data_ptr(),
                                   flow.rho0_shift, mask_ptr, out_ws.data_ptr(), ws.data_ptr(),
                                   stream), "lbm_reduce_f64")
    N.check(N.lib().lbm_reduce_f64_static(S.native_id, STORAGE_OF[f.dtype], n3, s4,
                                          f.data_ptr(), flow.rho0_shift, mask_ptr,
                                          out_st.data_ptr(), stream), "lbm_reduce_f64_static")
    torch.cuda.synchronize()
    return out_ws.cpu(), out_st.cpu()


@pytest.mark.parametrize("storage", ["f32", "bf16_shifted"])
def test_static_reduction_matches_the_workspace_reduction(storage):
    ctx = lt.Context("cuda:0", dtype=torch.float32, storage=storage)
    big = lt.TaylorGreenVortex(ctx, 96, 100, 0.1, lt.D3Q19(), initialize_fneq=False)
    small = lt.TaylorGreenVortex(ctx, 24, 100, 0.1, lt.D3Q19(), initialize_fneq=False)
    lt.Simulation(small, lt.BGKCollision(0.6), [], lt.StreamingStrategy.PRE_STREAMING)(3)
    for flow in (big, small):        # the second call on a grid with fewer partials
        ref, got = _both(flow)
        assert torch.equal(got, ref), (got, ref)
        assert float(ref[0]) > 0 and float(ref[4]) > 0
    mask = torch.zeros(small.resolution, dtype=torch.bool)
    mask[:, :5] = True
    ref, got = _both(small, mask)
    assert torch.equal(got, ref)
    full, _ = _both(small)
    assert float(ref[0]) < float(full[0])


def test_static_reduction_2d_after_3d():
    ctx = lt.Context("cuda:0", dtype=torch.float64)
    flow = lt.TaylorGreenVortex(ctx, 8, 100, 0.1, lt.D2Q9(), initialize_fneq=False)
    ref, got = _both(flow)
    assert torch.equal(got, ref)
    assert abs(float(got[0]) - 64.0) <= 1e-12 * 64
