"""The allocation-free TV + Adam entry point (ernn_tv_adam_step_static)
against the workspace one (ernn_tv_adam_step), both through the C-ABI on the
same inputs: bit for bit, on two calls in a row, the second smaller, so the
second call runs over what the first left in the static scratch."""
import ctypes as C

import pytest
import torch

from ernn import _lib

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _desc(nx, ny, nf, step, log_params):
    d = _lib.TvAdamDesc()
    d.nx, d.ny, d.n_fields, d.log_params, d.step = nx, ny, nf, int(log_params), step
    d.lambda_tv, d.lr, d.beta1, d.beta2, d.eps, d.weight_decay = 3e-2, 1e-2, 0.9, 0.999, 1e-8, 1e-4
    for f in range(nf):
        d.lo[f], d.hi[f] = (0.5, 4.0) if log_params else (-2.0, 2.0)
    return d


def _state(nx, ny, nf, seed):
    g = torch.Generator().manual_seed(seed)
    mk = lambda s=1.0: [(s * torch.randn(nx, ny, generator=g)).to(DEV) for _ in range(nf)]
    return dict(theta=mk(), grad=mk(1e-2), m=mk(1e-3), v=[t.abs() for t in mk(1e-5)],
                tv=torch.zeros(nf, dtype=torch.float64, device=DEV))


def _clone(st):
    return {k: ([t.clone() for t in v] if isinstance(v, list) else v.clone()) for k, v in st.items()}


def _call(st, ws):
    c = _lib.TvAdamCall()
    for f in range(len(st["theta"])):
        c.theta[f], c.grad[f] = st["theta"][f].data_ptr(), st["grad"][f].data_ptr()
        c.m[f], c.v[f] = st["m"][f].data_ptr(), st["v"][f].data_ptr()
    c.tv = st["tv"].data_ptr()
    c.workspace = ws.data_ptr() if ws is not None else None
    return c


def _run(fn, d, st, ws):
    print(f"[no-alloc] {fn} nx={d.nx} ny={d.ny} fields={d.n_fields} step={d.step}", flush=True)
    rc = getattr(_lib.lib(), fn)(C.byref(d), C.byref(_call(st, ws)), C.c_void_p(_lib.stream_handle(DEV)))
    assert rc == 0, _lib.lib().ernn_last_error().decode()
    torch.cuda.synchronize()


def _same(a, b):
    for k in a:
        xs, ys = (a[k], b[k]) if isinstance(a[k], list) else ([a[k]], [b[k]])
        for x, y in zip(xs, ys):
            assert torch.equal(x, y), k


@pytest.mark.parametrize("log_params", [True, False])
def test_static_scratch_entry_point_matches_workspace_entry_point(log_params):
    # 301 columns: the one-cell TV kernel; 96 columns: the four-cell kernel
    for (nx, ny, nf, seed) in [(301, 301, 6, 1), (128, 96, 3, 2)]:
        ref = _state(nx, ny, nf, seed)
        got = _clone(ref)
        d0 = _desc(nx, ny, nf, 1, log_params)
        nbytes = C.c_int64()
        assert _lib.lib().ernn_tv_adam_workspace_bytes(C.byref(d0), C.byref(nbytes)) == 0
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
        for step in (1, 2):
            d = _desc(nx, ny, nf, step, log_params)
            _run("ernn_tv_adam_step", d, ref, ws)
            _run("ernn_tv_adam_step_static", d, got, None)
            _same(ref, got)
        assert torch.all(torch.isfinite(got["tv"])) and float(got["tv"].min()) > 0
