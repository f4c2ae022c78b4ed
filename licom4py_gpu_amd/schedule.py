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


def _as_tensor(x):
    """A torch view of a device array exported through DLPack (no copy)."""
    import torch

    return x if isinstance(x, torch.Tensor) else torch.from_dlpack(x)


def make_schedule_core(rk_type_barotr: int = 2, px: int = 1, py: int = 1, from_dlpack=None):
    """schedule.py:130-192: rk_type_barotr == 2 selects RK2, anything else RK3.

    The returned `schedule_core` takes torch tensors.  Given `from_dlpack`
    (the caller's array library's importer, e.g. ``jax.dlpack.from_dlpack``),
    it also takes any device arrays that export DLPack: the state and consts
    are viewed as tensors without a copy (the same buffers map to the same
    cached context on every call), and every returned field is handed back
    through `from_dlpack`, so a caller such as ``Schedule.run`` keeps its own
    array type (its ``jax.device_get`` snapshots keep working)."""
    rk_core = barotr_rk2_core if rk_type_barotr == 2 else barotr_rk3_core

    def schedule_core(state, consts, chunk_size, nbb, dtb):
        barotr_consts = tuple(_as_tensor(c) for c in consts[:N_BAROTR_CONSTS])
        s = tuple(_as_tensor(x) for x in state)
        for _ in range(int(chunk_size)):
            bs = rk_core(s[:N_BAROTR_STATE], barotr_consts, nbb, dtb, px, py)
            s = tuple(bs) + s[N_BAROTR_STATE:]
        if from_dlpack is not None:
            s = tuple(from_dlpack(x) for x in s)
        return s

    return schedule_core


# ---------------------------------------------------------------------------
# Chunked driver with asynchronous host snapshots (mymodule/schedule.py:246-313,
# mymodule/diag.py:45-75)
# ---------------------------------------------------------------------------
def global_diag(h0, ub, vb, halo: int = 3) -> dict:
    """max / min / mean of the interior of h0, ub, vb (diag.py:45-64), from host
    or device arrays."""
    import numpy as np

    out = {}
    for name, a in (("h0", h0), ("ub", ub), ("vb", vb)):
        a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
        x = a[..., halo:-halo, halo:-halo]
        out[name] = (float(np.max(x)), float(np.min(x)), float(np.mean(x)))
    return out


def format_diag(tag: str, stats: dict) -> str:
    """The report layout of diag.py:66-75."""
    lines = [tag, "                  max                  min                 mean"]
    for name in ("h0", "ub", "vb"):
        mx, mn, me = stats[name]
        lines.append(f"Global {name}: {mx:20.16f}  {mn:20.16f}  {me:20.16f}")
    lines.append("------------------------------------------------")
    return "\n".join(lines)


class Driver:
    """Runs `schedule_core` in chunks of baroclinic steps and snapshots (h0, ub,
    vb) to pinned host memory after each chunk without stopping the device:
    the copies run on a side stream (default priority) ordered after the
    chunk by an event, and a worker thread waits for them and calls
    `on_snapshot(tag, h0, ub, vb)` with numpy arrays (the default prints the
    reference's diagnostics).  With ``donate`` the chunks update the state in
    place, and the next chunk waits for the previous snapshot's copies."""

    def __init__(self, state, consts, nbb: int, dtb: float, rk_type_barotr: int = 2,
                 donate: bool = False, on_snapshot=None):
        import concurrent.futures

        import torch

        self.state = tuple(state)
        self.consts = tuple(consts)
        self.nbb, self.dtb = int(nbb), float(dtb)
        self.donate = donate
        self.rk_core = barotr_rk2_core if rk_type_barotr == 2 else barotr_rk3_core
        self.device = self.state[0].device
        self.side = torch.cuda.Stream(self.device)
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        self.futures = []
        self.on_snapshot = on_snapshot or (lambda tag, h, u, v: print(format_diag(tag, global_diag(h, u, v))))
        self._last_copy = None

    def snapshot(self, tag: str):
        import torch

        src = (self.state[0], self.state[2], self.state[3])
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        host = [torch.empty(x.shape, dtype=x.dtype, pin_memory=True) for x in src]
        with torch.cuda.stream(self.side):
            self.side.wait_event(ready)
            for d, s in zip(host, src):
                d.copy_(s, non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.side)
        self._last_copy = done

        def finish(done=done, host=host, src=src):  # src keeps the device tensors alive
            done.synchronize()
            self.on_snapshot(tag, *(x.numpy() for x in host))

        self.futures.append(self.pool.submit(finish))

    def _chunk(self, n: int):
        import torch

        if self.donate and self._last_copy is not None:
            torch.cuda.current_stream(self.device).wait_event(self._last_copy)
        s = self.state
        for _ in range(n):
            bs = self.rk_core(s[:N_BAROTR_STATE], self.consts[:N_BAROTR_CONSTS], self.nbb, self.dtb,
                              donate=self.donate)
            s = tuple(bs) + s[N_BAROTR_STATE:]
        self.state = s

    def run(self, total_steps: int, diag_interval=None):
        """`total_steps` baroclinic steps (each nbb barotropic substeps), a
        snapshot at the start, after every `diag_interval` steps (if given)
        and at the end (schedule.py:272-313); returns the final state after all
        snapshots have been handled."""
        chunk = int(diag_interval) if diag_interval else int(total_steps)
        self.snapshot("Initial time")
        done = 0
        while done < total_steps:
            n = min(chunk, total_steps - done)
            self._chunk(n)
            done += n
            if diag_interval:
                self.snapshot(f"Time: step {done}")
        if not diag_interval:
            self.snapshot("Final time")
        for f in self.futures:
            f.result()
        self.futures.clear()
        return self.state
