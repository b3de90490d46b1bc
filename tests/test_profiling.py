"""The rocprofv3 summary (film_grain_amd.profiling) on small synthetic result
databases with the views it reads: kernel times, the copy calibration of
FETCH_SIZE / WRITE_SIZE, scaled traffic, and the counter table."""
import sqlite3

import pytest

from film_grain_amd import profiling


def _db(path, times=(), counters=()):
    path.mkdir()
    con = sqlite3.connect(path / "run_results.db")
    con.execute("create table top_kernels (name text, total_calls int, average real)")
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real, dispatch_id int)")
    con.executemany("insert into top_kernels values (?, ?, ?)", times)
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", counters)
    con.commit()
    con.close()
    return str(path)


def test_traffic_is_scaled_by_the_copy_calibration(tmp_path):
    copy = "__amd_rocclr_copyBuffer"
    trace = _db(tmp_path / "t", times=[("splat_hbm_kernel<512>", 4, 5000.0), ("bin_count_kernel", 4, 1000.0)])
    # a 256 MiB copy reads FETCH = 0.5 x and WRITE = 1.0 x its KiB
    fetch = _db(tmp_path / "f", counters=[(copy, "FETCH_SIZE", 0.5 * 256 * 1024, 0),
                                          ("splat_hbm_kernel<512>", "FETCH_SIZE", 1000.0, 1),
                                          ("splat_hbm_kernel<512>", "FETCH_SIZE", 3000.0, 2)])
    write = _db(tmp_path / "w", counters=[(copy, "WRITE_SIZE", 256 * 1024, 0),
                                          ("splat_hbm_kernel<512>", "WRITE_SIZE", 100.0, 1)])
    pmc = _db(tmp_path / "p", counters=[("splat_hbm_kernel<512>", "SQ_INSTS_VALU", 10.0, 1),
                                        ("splat_hbm_kernel<512>", "SQ_INSTS_VALU", 30.0, 2)])
    s = profiling.summarize(trace, fetch, write, 256, "splat_hbm", [pmc])
    assert s["fetch_factor"] == pytest.approx(0.5) and s["write_factor"] == pytest.approx(1.0)
    k = s["kernel"]
    assert k["avg_us"] == 5000.0 and k["calls"] == 4
    assert k["traffic_bytes"] == pytest.approx((2000.0 / 0.5 + 100.0) * 1024)
    assert s["counters"]["splat_hbm_kernel<512>"]["SQ_INSTS_VALU"] == pytest.approx(20.0)
    md = profiling.markdown(s, "t", ["cmd"], ["note"])
    assert "SQ_INSTS_VALU" in md and "| `splat_hbm_kernel<512>` | 4 | 5000.0 |" in md


def test_trace_alone_and_missing_calibration(tmp_path):
    trace = _db(tmp_path / "t", times=[("k", 1, 2.0)])
    s = profiling.summarize(trace, None, None, 256, "k")
    assert s["kernel"]["traffic_bytes"] is None and s["fetch_factor"] is None
    assert "Counter calibration" not in profiling.markdown(s, "t", [], [])
    empty = _db(tmp_path / "f", counters=[("k", "FETCH_SIZE", 1.0, 0)])
    with pytest.raises(ValueError):
        profiling.summarize(trace, empty, empty, 256, "k")
