"""tatva_valu_calibration, the measurement kernel behind the VALU roofline
(include/tatva_amd.h): argument checks on the host, and on the GPU every
instruction form's chains against a float32 restatement of the same sequence."""

import numpy as np
import pytest

from tatva_amd import _native as nat

A, B = np.float32(0.9999), np.float32(1e-4)


def _expected(op, blocks, iters):
    t = np.arange(256, dtype=np.float32)
    at = A * (np.float32(1.0) + np.float32(1e-7) * t)
    bt = B * (np.float32(1.0) - np.float32(1e-7) * t)
    total = np.zeros(256, dtype=np.float32)
    for k in range(8):
        v = (np.float32(1e-3) * t + np.float32(k)).astype(np.float32)
        for _ in range(4 * iters):   # 32 instructions a trip over 8 chains
            if op == 0:      # fma(v, a, b); float64 holds v * a exactly
                v = (v.astype(np.float64) * float(A) + float(B)).astype(np.float32)
            elif op == 1:    # fma(at, bt, v)
                v = (at.astype(np.float64) * bt + v).astype(np.float32)
            elif op == 2:
                v = (v + bt).astype(np.float32)
            elif op == 3:
                v = (v * at).astype(np.float32)
            elif op == 4:    # fma(v, at, bt)
                v = (v.astype(np.float64) * at + bt).astype(np.float32)
            else:            # fma(a, bt, v)
                v = (float(A) * bt.astype(np.float64) + v).astype(np.float32)
        total = (total + v).astype(np.float32)
    return np.tile(total, blocks)


def test_calibration_rejects_bad_arguments():
    L = nat.lib()
    dummy = 256  # never dereferenced: every case fails its checks first
    for args in ((6, 1, 1, dummy), (-1, 1, 1, dummy), (0, 0, 1, dummy), (0, 1, 0, dummy),
                 (0, 1, 1, 0)):
        assert L.tatva_valu_calibration(*args, None) == nat.EINVAL, args


@pytest.mark.gpu
@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 5])
def test_calibration_chains_match_float32_restatement(op):
    import torch

    L = nat.lib()
    blocks, iters = 3, 16
    out = torch.full((blocks * 256,), float("nan"), dtype=torch.float32, device="cuda")
    nat.check(L.tatva_valu_calibration(op, blocks, iters, nat.dptr(out), nat.stream()))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    # the device may contract 1 + 1e-7 t into one rounding: a_t can differ by an ulp,
    # which the 64 products of op 3 grow to ~4e-6; one trip more or less moves every
    # form by 1e-4 or more
    np.testing.assert_allclose(got, _expected(op, blocks, iters), rtol=2e-5, atol=0)
