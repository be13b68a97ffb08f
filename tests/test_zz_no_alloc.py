"""psim_mlp_forward_noalloc (include/psim_surrogate.h) against psim_mlp_forward.

The no-allocation entry point runs the same kernels with the same launch shapes on buffers carved
from the library's static scratch array, so its output must equal psim_mlp_forward's bit for bit:
on two calls in a row, the second with fewer rows (the scratch keeps the first call's contents),
and a call beyond the array's capacity must be refused with a message.
"""
import ctypes

import numpy as np
import pytest

from psim_amd import surrogate

pytestmark = pytest.mark.gpu

_F = ctypes.POINTER(ctypes.c_float)


def _noalloc(net, x):
    L = surrogate._bind()
    f = L.psim_mlp_forward_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, _F, ctypes.c_int32, _F]
    x = np.ascontiguousarray(x, np.float32)
    y = np.full(x.shape[0], np.nan, np.float32)
    rc = f(net._h, x.ctypes.data_as(_F), int(x.shape[0]), y.ctypes.data_as(_F))
    return rc, y


def test_noalloc_forward_equals_forward():
    net = surrogate.Surrogate(seed=3)
    rng = np.random.default_rng(11)
    for rows in (5000, 777):
        print(f"no-alloc forward, {rows} rows", flush=True)
        x = rng.uniform(-1.0, 1.0, size=(rows, surrogate.N_FEATURES)).astype(np.float32)
        ref = net.predict(x)
        rc, y = _noalloc(net, x)
        assert rc == 0, net._L.psim_mlp_last_error(net._h).decode()
        np.testing.assert_array_equal(y, ref)
    net.close()


def test_noalloc_forward_takes_its_stated_maximum_and_refuses_more():
    """29056 rows fit the 64 MiB array for 32 -> 512 x 4 (include/psim_surrogate.h); 29057 do not."""
    net = surrogate.Surrogate(seed=3)
    x = np.random.default_rng(5).uniform(-1.0, 1.0, size=(29057, surrogate.N_FEATURES)).astype(np.float32)
    rc, y = _noalloc(net, x[:29056])
    assert rc == 0, net._L.psim_mlp_last_error(net._h).decode()
    np.testing.assert_array_equal(y, net.predict(x[:29056]))
    rc, _ = _noalloc(net, x)
    assert rc != 0
    assert "scratch" in net._L.psim_mlp_last_error(net._h).decode()
    net.close()
