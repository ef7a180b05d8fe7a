"""The bench.py JSON line: the fields a driver reads, on a small wedge.

The GPU case runs bench.main() in this process on a 480 x 160 wedge (no child
process: the reference's CPU sample is off, the CPU port times one
iteration).  The CPU case checks that more than one GPU is refused before
anything touches a device.
"""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _run(monkeypatch, capsys, args):
    monkeypatch.setattr(sys, "argv", ["bench.py"] + args)
    bench.main()
    return capsys.readouterr()


def test_more_than_one_gpu_is_refused(monkeypatch, capsys):
    with pytest.raises(SystemExit) as e:
        _run(monkeypatch, capsys, ["--gpus", "2"])
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err


def test_alg_bytes_follow_the_scope_model():
    """Per stage launch: N (nv b q read + nv dim b gradient read + nv b w0 read
    + nv b q write + (1 + dim) 8 V, x) + F ((dim + 1) 8 n, A + 2 * 4 c0, c1):
    1.219 GB on the 3840 x 1280 wedge in FP64."""
    N, F = 3840 * 1280, 9835520
    assert bench.alg_bytes("stage", N, 0, F, 0, 2, 8) == N * 184 + F * 32 == 1219133440


@pytest.mark.gpu
def test_bench_line_on_a_small_wedge(monkeypatch, capsys):
    out = _run(monkeypatch, capsys, ["--nx", "480", "--ny", "160", "--steps", "5", "--warmup", "2",
                                     "--ref-iters", "0", "--cpu-iters", "1"])
    line = json.loads(out.out.strip().splitlines()[-1])
    assert line["metric"] == bench.METRIC and line["unit"] == "cell-updates/s"
    assert line["n_gpus"] == 1 and line["steps"] == 5 and line["warmup"] == 2
    assert line["value"] > 0 and line["ms_per_step"] > 0 and line["higher_is_better"] is True
    assert line["dtype"] == "f64" and line["vs_baseline"] is None
    assert line["config"]["cells"] == 480 * 160
    r = line["roofline"]
    assert r["kernel"] == "stage" and r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert 0 < r["frac"] < 1 and abs(r["achieved"] / r["peak"] - r["frac"]) < 1e-3
    # kernel means over a replay of the timed region; traffic is recorded for the 4.9M-cell wedge only
    assert "replay" in r["timed_over"] and r["traffic"] is None
    # mean_ms is printed to 1e-4 ms, a few parts per thousand of this small wedge's launches
    assert r["achieved"] == pytest.approx(r["alg_bytes_per_launch"] / (r["mean_ms"] * 1e-3) / 1e9, rel=1e-2)
    cb = line["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] == 1 and cb["value"] > 0
    assert line["fp32"]["value"] > 0
