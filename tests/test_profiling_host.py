"""Host logic of licom4py_gpu_amd.profiling on small sqlite databases laid
out like rocprofv3's (a `kernels` view with name / start / duration, a
`top_kernels` summary and `pmc_events` rows), so the figures the committed
profiles and bench.py take from them are checked without a GPU."""

import json
import sqlite3

import pytest

from licom4py_gpu_amd import profiling


def _trace_db(path, durations_ns):
    db = sqlite3.connect(path)
    db.execute("create table kernels (name text, start integer, duration integer)")
    for k, (name, d) in enumerate(durations_ns):
        db.execute("insert into kernels values (?, ?, ?)", (name, 1000 * k, d))
    db.commit()
    db.close()


def test_dispatch_phases_split_last_before_rest(tmp_path):
    path = str(tmp_path / "t.db")
    # 3 early dispatches (1 us), 4 "timed" (2 us), 2 "graph-replayed" (5 us), other kernels interleaved
    seq = [("k_fused<double>", 1000)] * 3 + [("k_frame<double>", 9000)] + \
          [("k_fused<double>", 2000)] * 4 + [("k_fused<double>", 5000)] * 2
    _trace_db(path, seq)
    ph = profiling.dispatch_phases(path, "k_fused", 2, 4)
    assert ph["all"] == (9, pytest.approx((3 * 1 + 4 * 2 + 2 * 5) / 9))
    assert ph["last"] == (2, pytest.approx(5.0))
    assert ph["before"] == (4, pytest.approx(2.0))
    assert ph["others"] == (3, pytest.approx(1.0))
    # asking for more than there is clips instead of overlapping
    ph = profiling.dispatch_phases(path, "k_fused", 2, 100)
    assert ph["before"][0] == 7 and ph["others"][0] == 0
    assert profiling.dispatch_phases(path, "k_nothing", 2) is None


def test_traffic_scales_counters_by_the_copy(tmp_path):
    def pmc_db(path, rows):
        db = sqlite3.connect(path)
        db.execute("create table pmc_events (name text, counter_name text, dispatch_id integer, value real)")
        db.executemany("insert into pmc_events values (?, ?, ?, ?)", rows)
        db.commit()
        db.close()

    fetch, write = str(tmp_path / "f.db"), str(tmp_path / "w.db")
    # the copy of 1 MiB reads as 512 KiB of FETCH_SIZE (128 B requests counted as 64 B) and 1 MiB written;
    # the kernel's two dispatches (summed over two counter dimensions each) average 300 and 100 KiB
    pmc_db(fetch, [("MulFunctor", "FETCH_SIZE", 1, 512.0),
                   ("k_fused<double>", "FETCH_SIZE", 2, 100.0), ("k_fused<double>", "FETCH_SIZE", 2, 100.0),
                   ("k_fused<double>", "FETCH_SIZE", 3, 400.0)])
    pmc_db(write, [("MulFunctor", "WRITE_SIZE", 1, 1024.0),
                   ("k_fused<double>", "WRITE_SIZE", 2, 50.0), ("k_fused<double>", "WRITE_SIZE", 3, 150.0)])
    out = str(tmp_path / "t.json")
    assert profiling.main(["traffic", "--fetch", fetch, "--write", write, "--kernel", "k_fused",
                           "--copy-mib", "1", "--out", out]) == 0
    got = json.load(open(out))
    assert got["fetch_scale"] == pytest.approx(2.0) and got["write_scale"] == pytest.approx(1.0)
    assert got["fetch_kib_per_dispatch"] == pytest.approx(300.0)
    assert got["bytes_per_launch"] == pytest.approx((300.0 * 2 + 100.0) * 1024)


def test_one_round_band_height():
    """The tuner's extra candidate: the lowest band height whose fused-kernel
    waves fit one round (8 / 12 resident waves per CU), counted by brute force
    over heights, offered only beyond the fixed list (64 rows)."""
    from licom4py_gpu_amd.autotune import one_round_rows

    def brute(nx, fp64, cus):
        strips = -(-nx // 52)
        slots = (8 if fp64 else 12) * cus
        for rows in range(1, nx + 1):
            if 6 * strips * -(-nx // rows) <= slots:
                return rows if rows > 64 else 0
        return 0

    assert one_round_rows(1536, True, 256) == 140 and one_round_rows(1536, False, 256) == 91
    for nx in (7, 48, 96, 384, 768, 1000, 1536, 2048, 3072, 8192):
        for fp64 in (True, False):
            for cus in (80, 256, 304):
                assert one_round_rows(nx, fp64, cus) == brute(nx, fp64, cus), (nx, fp64, cus)
