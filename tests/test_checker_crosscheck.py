"""The device-side gate (harness.check_implementation, torch) against the
pinned numpy gate (oracle.check_implementation), on the CPU: both restate
reference/reference.py:137-188, and every GPU test relies on the first, so
the two must accept and reject the same outputs with the same messages and
the same scaled residual."""
import re

import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import harness


def _scaled(msg):
    m = re.search(r"scaled_reconstruction_residual=([0-9.eE+-]+)", msg)
    return float(m.group(1)) if m else None


def _variants(L):
    """Good factor and one planted defect per rejection branch."""
    out = {"good": L.copy()}
    b = L.copy(); b[0, -1, 0] = np.nan; out["nan"] = b
    b = L.copy(); b[-1, 0, L.shape[-1] - 1] = 0.25; out["upper"] = b
    b = L.copy(); b[0] = -b[0]; out["negdiag"] = b
    b = L.copy(); b[-1, -1, 0] += 0.05 * np.abs(L).max(); out["recon"] = b
    b = L.copy(); b[0] *= 1 + 1e-3; out["scaled"] = b
    return out


@pytest.mark.parametrize("case,cond", harness.FAMILIES)
@pytest.mark.parametrize("n", [1, 7, 32, 100])
def test_device_gate_matches_pinned_gate(oracle_mod, case, cond, n):
    a = harness.generate_input(3, n, cond, 600 + n, case, device="cpu")
    L64, info = oracle_mod.potrf64(a.numpy())
    assert (info == 0).all()
    for name, L in _variants(L64.astype(np.float32)).items():
        if n == 1 and name == "upper":
            continue
        ok_t, msg_t = harness.check_implementation(a, torch.from_numpy(L))
        ok_n, msg_n = oracle_mod.check_implementation(a.numpy(), L)
        assert ok_t == ok_n, (name, msg_t, msg_n)
        assert msg_t.split(":")[0].split("=")[0] == msg_n.split(":")[0].split("=")[0], (name, msg_t, msg_n)
        st, sn = _scaled(msg_t), _scaled(msg_n)
        if st is not None:
            # the two FP32 products L L^T round in different orders: they
            # agree to a fraction of the n*eps unit the value is scaled by
            assert abs(st - sn) <= 0.05 * sn + 0.25, (name, st, sn)


def test_device_gate_rejects_shape_dtype_device():
    a = harness.generate_input(2, 8, 2, 1, device="cpu")
    L = torch.linalg.cholesky(a)
    assert not harness.check_implementation(a, L[:1])[0]
    assert not harness.check_implementation(a, L.double())[0]
    assert harness.check_implementation(a, L)[0]
