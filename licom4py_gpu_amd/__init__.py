"""MI355X-native barotropic shallow-water step on the cubed-sphere duo-grid.

Public surface (mirrors the reference's barotropic path):
  barotr.barotr_rk2_core / barotr_rk3_core / step_rk_logic — the drop-in step
  schedule.make_schedule_core                              — schedule_core contract
  duogrid.generate / load_npz / barotr_consts              — grid tables, const tuple
  initial.seeded_balanced / w92_case2 / state_tuple        — initial states
and beyond the reference's single field per process:
  ensemble.Batch                                           — many fields per launch
  serving.Server                                           — continuous batching under an HBM budget
  precision                                                — FP32 vs FP64 report
The HIP library is loaded lazily on first use (import does not touch the GPU).
"""

from .barotr import (CONST_SOURCES, N_BAROTR_CONSTS, N_BAROTR_STATE, STATE_KEYS,  # noqa: F401
                     barotr_rk2_core, barotr_rk3_core, step_rk_logic)
from .schedule import make_schedule_core  # noqa: F401

__version__ = "0.1.0"
