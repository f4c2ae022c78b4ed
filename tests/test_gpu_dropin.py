"""The drop-in recipes as a caller holding its own array type uses them (GPU).

The reference keeps its state in JAX arrays and would reach the HIP step
through DLPack.  A caller that rewraps the same const buffers on every call
must hit the cached context (no repack, no new workspace), and
`make_schedule_core(..., from_dlpack=...)` must take and return the caller's
arrays.  A small wrapper type exporting DLPack stands in for JAX arrays
(jax is not installed here)."""

from __future__ import annotations

import numpy as np
import pytest

from tests._util import seeded_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


class Foreign:
    """A device array of another library: it exports DLPack and nothing else."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__(stream=stream)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_rewrapped_consts_build_one_context():
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(10)
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(consts, dev)
    dstate = tuple(torch.as_tensor(x).to(dev) for x in state)
    before = barotr.context_builds()
    outs = []
    for wrap in (lambda c: torch.from_dlpack(Foreign(c)), lambda c: c.view(c.shape), lambda c: c[...]):
        outs.append(barotr.barotr_rk2_core(dstate, tuple(wrap(c) for c in dconsts), 2, dt))
    assert barotr.context_builds() - before == 1
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert torch.equal(a, b)
    # a modified const is a different input: a new context
    d2 = list(dconsts)
    d2[4] = dconsts[4].clone()
    d2[4].add_(0.0)
    barotr.barotr_rk2_core(dstate, tuple(d2), 1, dt)
    assert barotr.context_builds() - before == 2


def test_schedule_core_takes_and_returns_dlpack_arrays():
    torch = _torch()
    from licom4py_gpu_amd import barotr
    from licom4py_gpu_amd.schedule import make_schedule_core
    state, consts, dt = seeded_state(12)
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(consts, dev)
    dstate = tuple(torch.as_tensor(x).to(dev) for x in state) + (torch.zeros(3, device=dev),)  # + a tracer slot
    core = make_schedule_core(2, from_dlpack=Foreign)
    out = core(tuple(Foreign(x) for x in dstate), tuple(Foreign(c) for c in dconsts), 2, 3, dt)
    assert len(out) == 24 and all(isinstance(x, Foreign) for x in out)
    want = dstate
    for _ in range(2):
        want = tuple(barotr.barotr_rk2_core(want[:23], dconsts, 3, dt)) + want[23:]
    for k, (a, b) in enumerate(zip(out, want)):
        assert torch.equal(torch.from_dlpack(a), b), k
    # the results match the oracle: the JAX-side caller gets the reference's state
    from oracle import barotr_np as B
    ref = state
    for _ in range(2):
        ref = B.barotr_rk2_core(ref, consts, 3, dt)
    assert np.array_equal(torch.from_dlpack(out[0]).cpu().numpy(), ref[0])
