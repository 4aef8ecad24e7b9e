"""tatva_dot_noalloc: the allocation-free variant of tatva_dot (its partial sums
in a static device array of the library instead of a caller workspace), called
through the C ABI: the same value as tatva_dot, bit for bit, on two calls in a
row, the second on a smaller input (the scratch keeps the first call's partial
sums; the second must not read them), and the same rejections."""

import numpy as np
import pytest

from tatva_amd import _native as nat

pytestmark = pytest.mark.gpu


def _dot_ws(L, n, x, y):
    import torch

    out = torch.empty((), dtype=torch.float64, device="cuda")
    ws = torch.empty(int(L.tatva_dot_workspace(n)), dtype=torch.uint8, device="cuda")
    nat.check(L.tatva_dot(n, nat.dptr(x), nat.dptr(y), nat.dptr(out), nat.dptr(ws), ws.numel(),
                          nat.stream()))
    return out


def test_dot_noalloc_matches_dot():
    import torch

    L = nat.lib()
    fn = L.tatva_dot_noalloc
    rng = np.random.default_rng(3)
    big = 5_000_003          # the grid's full 2048 blocks
    x = torch.as_tensor(rng.standard_normal(big), device="cuda")
    y = torch.as_tensor(rng.standard_normal(big), device="cuda")
    for n in (big, 70_001, 1):  # each call smaller than the one before
        out = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
        nat.check(fn(n, nat.dptr(x), nat.dptr(y), nat.dptr(out), nat.stream()))
        ref = _dot_ws(L, n, x, y)
        torch.cuda.synchronize()
        assert float(out) == float(ref), (n, float(out), float(ref))
        host = float(np.dot(x[:n].cpu().numpy(), y[:n].cpu().numpy()))
        assert abs(float(out) - host) <= 1e-14 * np.sqrt(n) * max(abs(host), 1.0)
    with pytest.raises(ValueError, match="invalid arguments"):
        nat.check(fn(0, nat.dptr(x), nat.dptr(y), nat.dptr(out), nat.stream()))
