"""lh_order0_charge_noalloc (fixed device scratch, no allocation) equals
lh_order0_charge (caller workspace) bit for bit, on two calls in a row, the
second with a smaller input; and a call beyond the scratch limit is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _charge_with_workspace(order0, elements):
    import torch
    out, ws = order0.charge_device(elements)
    torch.cuda.synchronize()
    return out.cpu().numpy(), int(ws[:4].view(dtype=torch.int32)[0].item())


def _charge_noalloc(order0, elements):
    import torch
    out = torch.full((order0.oms.count,), float("nan"), dtype=torch.float64, device="cuda")
    flags = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    order0.charge_noalloc_device(elements, out, flags)
    torch.cuda.synchronize()
    return out.cpu().numpy(), int(flags.item())


def test_noalloc_charge_equals_workspace_charge(device):
    from lighthit_amd.ballistic import DeviceElements, Order0
    from lighthit_amd.model import acceptance_alpha
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    order0 = Order0(w.medium, w.detector, acceptance_alpha(w.detector))
    large = DeviceElements(w.elements)
    small = DeviceElements(w.elements.subset(np.arange(len(w.elements)) % 5 == 0))
    print("noalloc: demonstration event, 2000 then 400 elements", flush=True)
    for elements in (large, small):
        want, want_flags = _charge_with_workspace(order0, elements)
        got, got_flags = _charge_noalloc(order0, elements)
        assert got_flags == want_flags == 0
        assert np.array_equal(got, want), np.max(np.abs(got - want))
        assert want.sum() > 0


def test_noalloc_refuses_inputs_beyond_its_scratch(device):
    from lighthit_amd import _lib
    from lighthit_amd.ballistic import Order0, _Elements, _workspace_bytes
    from lighthit_amd.model import acceptance_alpha
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    order0 = Order0(w.medium, w.detector, acceptance_alpha(w.detector))
    fn = _lib.load().lh_order0_charge_noalloc_max_bytes
    fn.restype, fn.argtypes = ctypes.c_size_t, []
    limit = int(fn())
    # an element count whose workspace exceeds the scratch; the check runs on
    # the host before any device work, so the (null) element arrays are never read
    count = 1
    while _workspace_bytes(order0.oms.count, count) <= limit:
        count *= 2
    fake = _Elements(None, None, None, None, None, None, None, None, count)
    with pytest.raises(ValueError, match="scratch"):
        _lib.call("lh_order0_charge_noalloc", ctypes.byref(order0.optics.struct),
                  ctypes.byref(order0.oms.struct), ctypes.byref(fake), None, None,
                  _lib.stream_handle())
