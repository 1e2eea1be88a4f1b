"""The no-allocation element entry point (sparlab_elements_evaluate_noalloc_device)
against sparlab_elements_evaluate_device on the same inputs: the same kernels
on the library's scratch array instead of a caller workspace, so every output
must agree bit for bit. Two calls in a row, the second on a smaller batch (the
scratch keeps the first call's contents), and the refusal of a batch larger
than the scratch, which is made before any device call.
"""
import numpy as np
import pytest
import torch

import sparlab_gpu as G
from sparlab_gpu import workloads as W

OUTPUTS = ("force", "tangent", "alpha_out", "state_out", "energy", "yielding", "iterations", "status")


def batches(ne, seed, device):
    m = W.microbench_material()
    w = W.microbench(ne, seed=seed)
    st = G.states_from_reference(w["states"], m)
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    inputs = (t(w["coords"]), t(w["conn"], torch.int32), t(w["u"]), t(st))
    return m, G.ElementBatch(*inputs, None, 8), G.ElementBatch(*inputs, None, 8)


@pytest.mark.gpu
def test_noalloc_matches_the_workspace_entry_point(device):
    rule = G.Rule()
    for ne, seed in ((6000, 5), (700, 6)):       # the second call smaller than the first
        print(f"no-alloc parity: {ne} microbenchmark elements, seed {seed}", flush=True)
        m, ref, out = batches(ne, seed, device)
        for name in OUTPUTS:                       # outputs must come from the call, not from the buffers
            getattr(out, name).fill_(7)
        G.evaluate_elements(m, rule, ref)
        G.evaluate_elements_noalloc(m, rule, out)
        torch.cuda.synchronize()
        assert int((ref.status != 0).sum()) == 0
        assert int(ref.yielding.sum()) > 0
        for name in OUTPUTS:
            a, b = getattr(ref, name), getattr(out, name)
            assert torch.equal(a, b), f"{name} differs between the two entry points ({ne} elements)"


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    """Refused on the host before any device call (so it runs without a GPU):
    the pointers are never dereferenced."""
    import ctypes
    m = W.microbench_material()
    rule = G.Rule()
    per_element = G.workspace_doubles(2, 8) - G.workspace_doubles(1, 8)
    ne = G.NOALLOC_SCRATCH_BYTES // (8 * per_element) + 1000
    assert 8 * G.workspace_doubles(ne, 8) > G.NOALLOC_SCRATCH_BYTES
    fake = ctypes.c_void_p(4096)
    b = G.CElementBatch(fake, fake, fake, 8 * ne, ne, fake, None, fake, fake, fake, fake, fake, fake, fake, fake,
                        fake, None)
    code = G.lib().sparlab_elements_evaluate_noalloc_device(ctypes.byref(m.c), ctypes.byref(rule.c), 2, 1,
                                                            ctypes.byref(b), None)
    assert code == -1
    assert "scratch" in G.lib().sparlab_last_error().decode()
