"""The live HBM-traffic figure of bench.py: parsing of rocprofv3 counter rows and the calibration
on the 8-byte scale kernel of the bandwidth probe (CPU only; the child runs need the GPU)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


F64 = "void hc::k_stage<1, 1, 5, double, 7, 16, 256, 1, 1, 0>(hc::LaunchArgs<double>)"
F64_TUNE = "void hc::k_stage<1, 1, 5, double, 4, 32, 256, 0, 0, 0>(hc::LaunchArgs<double>)"
F32 = "void hc::k_stage<1, 1, 5, float, 7, 16, 256, 0, 1, 0>(hc::LaunchArgs<float>)"
UPD = "void hc::k_stage0_update<1, 1, double>(hc::LaunchArgs<double>, hc::UpdChunk const*)"
SCALE = "(anonymous namespace)::k_scale8(double*, double const*, double, long long)"


def _rows(counter, table):
    return [{"Kernel_Name": k, "Counter_Name": counter, "Counter_Value": str(v)} for k, vs in table for v in vs]


def test_stage_counter_picks_the_most_dispatched_stage_kernel_of_the_precision():
    b = _bench()
    n = 1 << 27   # 1 GiB each way
    rows = _rows("FETCH_SIZE", [(F64, [100.0, 110.0, 120.0]), (F64_TUNE, [999.0]), (F32, [5.0] * 9),
                                (UPD, [7.0] * 20), (SCALE, [524288.0])])
    rows += _rows("WRITE_SIZE", [(F64, [1.0e9])])   # another counter is ignored
    name, kib, scale = b.stage_counter(rows, "FETCH_SIZE", "double", n)
    assert name == F64
    assert kib == pytest.approx(110.0)
    assert scale == pytest.approx(2.0)   # the scale kernel reads half of its 1 GiB in this counter
    name, kib, _ = b.stage_counter(rows, "FETCH_SIZE", "float", n)
    assert name == F32 and kib == pytest.approx(5.0)


def test_stage_counter_needs_both_kernels():
    b = _bench()
    assert b.stage_counter(_rows("FETCH_SIZE", [(F64, [1.0])]), "FETCH_SIZE", "double", 1 << 20) is None
    assert b.stage_counter(_rows("FETCH_SIZE", [(SCALE, [1.0])]), "FETCH_SIZE", "double", 1 << 20) is None
    assert b.stage_counter(_rows("FETCH_SIZE", [(F64, [1.0]), (SCALE, [0.0])]), "FETCH_SIZE", "double",
                           1 << 20) is None


def test_traffic_entry_falls_back_to_the_committed_figure(monkeypatch):
    b = _bench()

    class A:
        traffic = "live"

    monkeypatch.setattr(b, "traffic_live", lambda a, p, t, n: (None, "rocprofv3 not found"))
    v, info = b.traffic_entry(A(), "c4", 0, 2, False, 1 << 27)
    assert v == b.traffic_from_profiles("c4", 0, 2) and v > 0
    assert info["source"] == "profiles/traffic.json" and info["live_error"] == "rocprofv3 not found"
    monkeypatch.setattr(b, "traffic_live", lambda a, p, t, n: (123.0, {"kernel": F64}))
    v, info = b.traffic_entry(A(), "c4", 0, 2, False, 1 << 27)
    assert v == 123.0 and info["committed_profile_bytes"] == b.traffic_from_profiles("c4", 0, 2)


def test_launcher_replaces_the_childs_traffic_with_the_live_figure(monkeypatch, capsys):
    import json
    import subprocess
    import sys
    b = _bench()
    line = {"metric": "m", "value": 1.0, "config": {"precision": "fp64", "tile_shape": 2, "cells": 1000},
            "roofline": {"traffic": 5.0}}
    calls = []

    def fake_run(cmd, **kw):
        calls.append(cmd)
        return subprocess.CompletedProcess(cmd, 0, stdout="# progress\n" + json.dumps(line) + "\n")

    monkeypatch.setattr(subprocess, "run", fake_run)
    monkeypatch.setattr(b, "traffic_live", lambda a, p, t, n: (2000.0, {"kernel": F64}))
    monkeypatch.setattr(sys, "argv", ["bench.py", "--workload", "c4"])

    class A:
        traffic = "live"
        workload = "c4"

    b.launch_with_live_traffic(A())
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert calls and calls[0][-2:] == ["--traffic", "file"]   # the child keeps the committed figure
    assert out["roofline"]["traffic"] == 2000.0
    assert out["roofline"]["traffic_bytes_per_cell_stage"] == 2.0
    assert out["roofline"]["traffic_measurement"]["kernel"] == F64


def test_launcher_passes_a_failed_measurement_through(monkeypatch, capsys):
    import subprocess
    import sys
    b = _bench()
    monkeypatch.setattr(subprocess, "run", lambda cmd, **kw: subprocess.CompletedProcess(cmd, 3, stdout=""))
    monkeypatch.setattr(sys, "argv", ["bench.py"])

    class A:
        traffic = "live"
        workload = "c4"

    with pytest.raises(SystemExit) as e:
        b.launch_with_live_traffic(A())
    assert e.value.code == 3
