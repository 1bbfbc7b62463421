"""abhsf_mlp_train_noalloc against abhsf_mlp_train, through the C ABI.

The no-alloc entry point runs the same kernels on a work buffer held in a
device array of the library instead of a per-call hipMalloc, so its trained
parameters and per-step losses must equal abhsf_mlp_train's bit for bit.  Two
calls in a row, the second on fewer samples and steps (and with sample
weights), check that nothing of the first call's buffer contents leaks into
the second.
"""
import ctypes as C

import numpy as np
import pytest

from abhsf_amd import capi, hip, tuner
from abhsf_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

F = 26  # the tuner's feature count


def problem(n, seed, weighted):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, F)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.2 * X[:, 3] ** 2).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32) if weighted else None
    return X, y, w


def train_both(n, steps, seed, weighted, lr=1e-2):
    X, y, w = problem(n, seed, weighted)
    theta0 = tuner.init_theta(F, seed).astype(np.float32)
    dX, dy = DeviceArray.from_numpy(X), DeviceArray.from_numpy(y)
    dw = None if w is None else DeviceArray.from_numpy(w)
    wp = None if dw is None else dw.ptr
    # reference: the allocating entry point (synchronous, losses to the host)
    t_ref = DeviceArray.from_numpy(theta0)
    loss_ref = np.zeros(steps, dtype=np.float32)
    capi.call("abhsf_mlp_train", t_ref.ptr, F, dX.ptr, dy.ptr, wp, n, steps, lr,
              loss_ref.ctypes.data_as(C.c_void_p), None)
    # no-alloc: asynchronous on the stream, losses to a device array
    st = hip.Stream()
    t_na = DeviceArray.from_numpy(theta0)
    loss_na = DeviceArray.zeros(steps, np.float32)
    capi.call("abhsf_mlp_train_noalloc", t_na.ptr, F, dX.ptr, dy.ptr, wp, n, steps, lr, loss_na.ptr, st.handle)
    st.synchronize()
    return t_ref.numpy(), loss_ref, t_na.numpy(), loss_na.numpy()


def test_noalloc_train_equals_train_twice(gpu):
    print("case noalloc train n=3000 steps=60, then n=700 steps=25 weighted", flush=True)
    for n, steps, seed, weighted in ((3000, 60, 1, False), (700, 25, 2, True)):
        th_ref, l_ref, th_na, l_na = train_both(n, steps, seed, weighted)
        assert np.array_equal(th_ref, th_na), (n, steps)
        assert np.array_equal(l_ref, l_na), (n, steps)
        assert l_na[-1] < l_na[0]  # it trained


def test_noalloc_train_without_loss_output(gpu):
    """loss may be NULL: the per-step losses then stay in the work buffer."""
    print("case noalloc train loss=NULL", flush=True)
    X, y, _ = problem(500, 3, False)
    theta0 = tuner.init_theta(F, 3).astype(np.float32)
    dX, dy = DeviceArray.from_numpy(X), DeviceArray.from_numpy(y)
    a, b = DeviceArray.from_numpy(theta0), DeviceArray.from_numpy(theta0)
    capi.call("abhsf_mlp_train", a.ptr, F, dX.ptr, dy.ptr, None, 500, 10, 1e-2, None, None)
    capi.call("abhsf_mlp_train_noalloc", b.ptr, F, dX.ptr, dy.ptr, None, 500, 10, 1e-2, None, None)
    hip.synchronize()
    assert np.array_equal(a.numpy(), b.numpy())


def test_noalloc_train_rejects_sizes_beyond_its_buffer(gpu):
    print("case noalloc train limits", flush=True)
    X, y, _ = problem(8, 4, False)
    dX, dy = DeviceArray.from_numpy(X), DeviceArray.from_numpy(y)
    t = DeviceArray.from_numpy(tuner.init_theta(F, 4).astype(np.float32))
    with pytest.raises(capi.AbhsfError, match="scratch limit of 4096 samples") as e:
        capi.call("abhsf_mlp_train_noalloc", t.ptr, F, dX.ptr, dy.ptr, None, 4097, 1, 1e-2, None, None)
    assert e.value.status == capi.ERR_UNSUPPORTED
    with pytest.raises(capi.AbhsfError, match="scratch limit of 16384 steps"):
        capi.call("abhsf_mlp_train_noalloc", t.ptr, F, dX.ptr, dy.ptr, None, 8, 16385, 1e-2, None, None)
