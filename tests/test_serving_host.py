"""Host-side logic of the serving loop (CPU): HBM accounting and the FCFS
comparison model."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from licom4py_gpu_amd import serving
from licom4py_gpu_amd.serving import CONST_PLANES, SCRATCH_FIELDS, STATE_FIELDS, HBMLedger, field_bytes


def test_field_bytes():
    assert field_bytes(48, torch.float64) == 6 * 54 * 54 * 8
    assert field_bytes(48, torch.float32) == 6 * 54 * 54 * 4


def test_ledger_first_admission_pays_consts_and_workspace():
    k = (48, torch.float64)
    L = HBMLedger(1e12)
    f = field_bytes(*k)
    assert L.need(k) == (SCRATCH_FIELDS + STATE_FIELDS + CONST_PLANES) * f
    assert L.admit(k) is False
    assert L.used == (SCRATCH_FIELDS + STATE_FIELDS + CONST_PLANES) * f
    # the consts are resident now
    assert L.need(k) == (SCRATCH_FIELDS + STATE_FIELDS) * f


def test_ledger_retire_keeps_workspace_idle_and_reuses_it():
    k = (96, torch.float32)
    f = field_bytes(*k)
    L = HBMLedger(1e12)
    L.admit(k)
    L.retire(k)
    assert L.used == (SCRATCH_FIELDS + CONST_PLANES) * f  # state freed, workspace idle
    assert L.need(k) == STATE_FIELDS * f
    assert L.admit(k) is True                               # reuses the idle workspace
    assert L.used == (SCRATCH_FIELDS + STATE_FIELDS + CONST_PLANES) * f
    assert L.idle[k] == 0


def test_ledger_budget_and_drop_idle():
    a, b = (48, torch.float64), (96, torch.float64)
    fa, fb = field_bytes(*a), field_bytes(*b)
    budget = (SCRATCH_FIELDS + STATE_FIELDS + CONST_PLANES) * (fa + fb)
    L = HBMLedger(budget)
    assert L.fits(a) and L.admit(a) is False
    assert L.fits(b)
    L.admit(b)
    assert not L.fits(b)               # a second C96 does not fit
    L.retire(a)
    assert L.drop_idle() == a          # the idle C48 workspace is released
    assert L.drop_idle() is None
    assert L.used == (SCRATCH_FIELDS + STATE_FIELDS + CONST_PLANES) * fb + CONST_PLANES * fa
    assert L.peak == budget


def test_drop_idle_takes_largest_first():
    L = HBMLedger(1e12)
    for k in ((48, torch.float64), (384, torch.float32), (96, torch.float64)):
        L.admit(k)
        L.retire(k)
    assert L.drop_idle() == (384, torch.float32)


def test_fcfs_latencies():
    lat, end = serving.fcfs_latencies([0.0, 0.5, 3.0], [1.0, 1.0, 1.0])
    assert np.allclose(lat, [1.0, 1.5, 1.0])
    assert end == 4.0


def test_latency_stats():
    reqs = [serving.Request(k, 8, 1, torch.float64, 1.0, (None, None, None), arrival=float(k),
                            t_admit=float(k), t_done=float(k) + 1 + k) for k in range(10)]
    st = serving.latency_stats(reqs)
    assert st["n"] == 10 and st["max_s"] == 10.0 and st["p50_s"] == pytest.approx(5.5)


def test_global_diag_and_format():
    from licom4py_gpu_amd.schedule import format_diag, global_diag
    h = np.zeros((6, 10, 10))
    h[:, 3:-3, 3:-3] = np.arange(6 * 16).reshape(6, 4, 4)
    h[0, 0, 0] = 1e9  # halo values are ignored
    st = global_diag(h, -h, h * 2)
    assert st["h0"] == (95.0, 0.0, 47.5)
    assert st["ub"] == (0.0, -95.0, -47.5)
    text = format_diag("Final time", st)
    assert text.splitlines()[0] == "Final time"
    assert "Global h0:" in text and "47.5000000000000000" in text
