"""The device math helpers' rounding on the GPU (tests/cpp/device_math_probe.hip).

recip_positive (v_rcp_f64 and two Newton-Raphson steps) and rsqrt_positive
(v_rsq_f64 and two Newton-Raphson steps) stand in for IEEE 1 / x and
1 / sqrt(x) in the general return and the log strain, on arguments whose
range is known (positive normal numbers). Over their domains (log-uniform
samples of [1e-280, 1e280] and [1e-290, 1e290], and the log strain's [1, 2])
each is within one ulp of the correctly rounded value (numpy's division;
1 / sqrt(x) from 50-digit decimal arithmetic, rounded by float()), and the
device's IEEE division matches numpy's bit for bit.
"""
import ctypes
import decimal
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "libdevice_math_probe.so")


def ulps(a, b):
    """|a - b| in units in the last place, for positive doubles."""
    return np.abs(a.view(np.int64) - b.view(np.int64))


def probe(x, device):
    import torch
    lib = ctypes.CDLL(PROBE)
    lib.device_math_probe.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p]
    d_x = torch.as_tensor(x, device=device)
    out = torch.empty((x.size, 4), dtype=torch.float64, device=device)
    s = torch.cuda.current_stream(device)
    assert lib.device_math_probe(ctypes.c_void_p(d_x.data_ptr()), x.size, ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(s.cuda_stream)) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_reciprocal_within_an_ulp(device):
    rng = np.random.default_rng(7)
    x = np.concatenate([10.0 ** rng.uniform(-280, 280, 20000), rng.uniform(1.0, 2.0, 20000),
                        [1.0, 2.0, 3.0, np.nextafter(1.0, 2.0), np.nextafter(2.0, 1.0), 1e-280, 1e280]])
    out = probe(x, device)
    exact = 1.0 / x
    assert np.array_equal(out[:, 1], exact)  # the device's IEEE division
    err = ulps(out[:, 0], exact)
    print(f"recip_positive: {np.mean(err == 0):.4f} correctly rounded, at most {err.max()} ulp", flush=True)
    assert err.max() <= 1


def test_inverse_square_root_within_an_ulp(device):
    rng = np.random.default_rng(8)
    x = np.concatenate([10.0 ** rng.uniform(-290, 290, 3000), rng.uniform(1.0, 2.0, 3000),
                        [1.0, 2.0, 4.0, np.nextafter(1.0, 2.0), np.nextafter(2.0, 1.0), 1e-290, 1e290]])
    out = probe(x, device)
    decimal.getcontext().prec = 50
    exact = np.array([float(1 / decimal.Decimal(float(v)).sqrt()) for v in x])
    err = ulps(out[:, 2], exact)
    print(f"rsqrt_positive: {np.mean(err == 0):.4f} correctly rounded, at most {err.max()} ulp", flush=True)
    assert err.max() <= 1
    assert ulps(out[:, 3], exact).max() <= 1  # the IEEE 1 / sqrt(x), two roundings
