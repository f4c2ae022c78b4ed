"""`schedule_core` contract (mymodule/schedule.py:130-192).

`make_schedule_core(rk_type_barotr)` returns
`schedule_core(state, consts, chunk_size, nbb, dtb)`, which runs `chunk_size`
baroclinic steps, each calling the barotropic core on `state[:23]` with
`consts[:17]`; fields beyond the first 23 pass through (the reference's
baroclinic and tracer slots are TODO stubs there too).  Launches are queued on
the current stream and the call returns without waiting, as the reference's
jitted call returns futures.
"""

from __future__ import annotations

from .barotr import (N_BAROTR_CONSTS, N_BAROTR_STATE, STATE_KEYS, CONST_SOURCES,
                     barotr_rk2_core, barotr_rk3_core)

__all__ = ["make_schedule_core", "STATE_KEYS", "CONST_SOURCES",
           "N_BAROTR_STATE", "N_BAROTR_CONSTS"]


def make_schedule_core(rk_type_barotr: int = 2, px: int = 1, py: int = 1):
    """schedule.py:130-192: rk_type_barotr == 2 selects RK2, anything else RK3."""
    rk_core = barotr_rk2_core if rk_type_barotr == 2 else barotr_rk3_core

    def schedule_core(state, consts, chunk_size, nbb, dtb):
        barotr_consts = tuple(consts[:N_BAROTR_CONSTS])
        s = tuple(state)
        for _ in range(int(chunk_size)):
            bs = rk_core(s[:N_BAROTR_STATE], barotr_consts, nbb, dtb, px, py)
            s = tuple(bs) + s[N_BAROTR_STATE:]
        return s

    return schedule_core
