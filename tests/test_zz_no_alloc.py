"""mc_point_loss_static (the static-workspace point loss, which allocates
nothing) against mc_point_loss with a caller-provided workspace, both called
through the C-ABI: bit-identical loss and gradient, on two calls in a row
whose second input is smaller than the first (the static workspace keeps the
first call's contents)."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MC_OK, MC_ERR_BAD_SHAPE = 0, 1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from microcosmos_amd import _lib
    lib = _lib.load()
    return lib, torch.device("cuda", 0)


def _cloud(seed, n, m):
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=(n, 2)) * 20.0 + 128.0).astype(np.float32)
    t = (rng.normal(size=(m, 2)) * 15.0 + 120.0).astype(np.float32)
    # exact ties: a few particles repeated among the targets and each other
    k = min(n, m, 8)
    t[:k] = p[:k]
    if n > 2:
        p[n - 1] = p[0]
    return p, t


def _ptr(x):
    return ctypes.c_void_p(x.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _with_workspace(lib, dev, kind, p, t, alpha, want_grad):
    n, m = p.shape[0], t.shape[0]
    ws = torch.empty(int(lib.mc_point_loss_workspace(n, m)), dtype=torch.uint8, device=dev)
    ws.fill_(0xFF)  # NaN-filled: the kernels must write every word they read
    loss = torch.full((), float("nan"), dtype=torch.float32, device=dev)
    grad = torch.full((n, 2), float("nan"), dtype=torch.float32, device=dev)
    st = lib.mc_point_loss(kind, _ptr(p), n, _ptr(t), m, alpha, _ptr(loss),
                           _ptr(grad) if want_grad else None, _ptr(ws), _stream(dev))
    assert st == MC_OK, lib.mc_last_error()
    return loss, grad


def _static(lib, dev, kind, p, t, alpha, want_grad):
    n, m = p.shape[0], t.shape[0]
    loss = torch.full((), float("nan"), dtype=torch.float32, device=dev)
    grad = torch.full((n, 2), float("nan"), dtype=torch.float32, device=dev)
    st = lib.mc_point_loss_static(kind, _ptr(p), n, _ptr(t), m, alpha, _ptr(loss),
                                  _ptr(grad) if want_grad else None, _stream(dev))
    assert st == MC_OK, lib.mc_last_error()
    return loss, grad


def _same(a, b):
    return np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("want_grad", [True, False])
def test_static_matches_workspace_twice(env, kind, want_grad):
    lib, dev = env
    alpha = 3.0
    for call, (n, m, seed) in enumerate(((1500, 900, 11), (200, 77, 12))):
        _say(f"kind {kind} grad {want_grad} call {call} n {n} m {m}")
        p_np, t_np = _cloud(seed, n, m)
        p = torch.as_tensor(p_np, device=dev)
        t = torch.as_tensor(t_np, device=dev)
        ls, gs = _static(lib, dev, kind, p, t, alpha, want_grad)
        lw, gw = _with_workspace(lib, dev, kind, p, t, alpha, want_grad)
        torch.cuda.synchronize()
        assert np.isfinite(float(ls.item()))
        assert _same(ls, lw), (float(ls.item()), float(lw.item()))
        if want_grad:
            assert torch.isfinite(gs).all()
            assert _same(gs, gw)


def test_static_at_the_maximum_size(env):
    lib, dev = env
    n, m = 65536, 4096
    _say(f"chamfer at the maximum particle count n {n} m {m}")
    p_np, t_np = _cloud(21, n, m)
    p = torch.as_tensor(p_np, device=dev)
    t = torch.as_tensor(t_np, device=dev)
    ls, gs = _static(lib, dev, 0, p, t, 1.0, True)
    lw, gw = _with_workspace(lib, dev, 0, p, t, 1.0, True)
    torch.cuda.synchronize()
    assert _same(ls, lw) and _same(gs, gw)
    # then a smaller call with swapped roles reuses the same array
    ls2, gs2 = _static(lib, dev, 0, t[:300], p[:5000], 1.0, True)
    lw2, gw2 = _with_workspace(lib, dev, 0, t[:300], p[:5000], 1.0, True)
    torch.cuda.synchronize()
    assert _same(ls2, lw2) and _same(gs2, gw2)


def test_static_rejects_sizes_beyond_the_maximum(env):
    lib, dev = env
    p = torch.zeros((4, 2), dtype=torch.float32, device=dev)
    loss = torch.zeros((), dtype=torch.float32, device=dev)
    for n, m in ((65537, 4), (4, 65537)):
        st = lib.mc_point_loss_static(0, _ptr(p), n, _ptr(p), m, 1.0, _ptr(loss), None,
                                      _stream(dev))
        assert st == MC_ERR_BAD_SHAPE
        assert b"static workspace" in lib.mc_last_error()
