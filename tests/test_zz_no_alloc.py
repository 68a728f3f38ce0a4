"""The allocation-free solve (amp_solve_device_noalloc) through the C ABI: same results as amp_solve_device on two
calls in a row, the second with a smaller input (another, smaller system and fewer right-hand sides), and requests
beyond its fixed scratch refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import ampsolve as amp
from ampsolve import _capi
from ampsolve import workloads as W

pytestmark = pytest.mark.gpu


def solve_both(device, ncell, q, kind):
    pat, v, _ = W.heat3d(ncell)
    n = pat["n"]
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, kind)
    assert num.factor(v)
    B = np.concatenate([2 * W.lcg_uniform(50 + k, n) - 1 for k in range(q)])
    dB = device.DeviceArray.from_numpy(B)
    X1 = device.DeviceArray.zeros(B.shape)
    X2 = device.DeviceArray.zeros(B.shape)
    _capi.check(_capi.lib().amp_solve_device(num._h, _capi.ptr(dB), _capi.ptr(X1), q))
    _capi.check(_capi.lib().amp_solve_device_noalloc(num._h, _capi.ptr(dB), _capi.ptr(X2), q))
    x1, x2 = X1.numpy(), X2.numpy()
    rel = np.abs(x1).max()
    print(f"noalloc: heat3d({ncell}) n={n} q={q} {kind}: max |x| {rel:.3e}", flush=True)
    return x1, x2, num, dB, X2


@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
def test_noalloc_solve_matches_the_allocating_solve(device, kind):
    # first call: the larger input (tiled supernodes, several passes of the flow kernels)
    x1, x2, num, dB, X2 = solve_both(device, 16, 11, kind)
    assert np.array_equal(x1, x2)  # same kernels, same order: bitwise
    # second call in a row: a smaller system and one right-hand side in the same scratch
    y1, y2, _, _, _ = solve_both(device, 8, 1, kind)
    assert np.array_equal(y1, y2)
    # beyond the fixed scratch: refused with an argument error, outputs untouched
    before = X2.numpy()
    rc = _capi.lib().amp_solve_device_noalloc(num._h, _capi.ptr(dB), _capi.ptr(X2), 100000)
    assert rc == _capi.AMP_ERR_ARG
    assert "scratch" in _capi.lib().amp_last_error().decode()
    assert np.array_equal(X2.numpy(), before)
