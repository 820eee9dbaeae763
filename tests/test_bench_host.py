"""bench.py's host logic on the CPU: the one-GPU refusal, the traffic
conversion and its source digest, and the neighbour-search figures."""
import importlib.util
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _csv(path, counter, rows):
    with open(path, "w") as f:
        f.write("Kernel_Name,Counter_Name,Counter_Value\n")
        for name, v in rows:
            f.write(f'"{name}",{counter},{v}\n')


def test_write_traffic_scales_by_the_copy(tmp_path, monkeypatch):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) scaled by what each reads
    for the calibration copy of --copy-mib MiB, averaged over dispatches;
    the file records the digest of the kernel sources, and a stale digest
    gives no traffic figure."""
    b = _bench()
    monkeypatch.setattr(b, "ROOT", tmp_path)
    (tmp_path / "profiles").mkdir()
    (tmp_path / "coupledwateranimation_amd" / "csrc").mkdir(parents=True)
    src = tmp_path / "coupledwateranimation_amd" / "csrc" / "k.hip"
    src.write_text("// kernel v1\n")
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    # the copy reads half its bytes on FETCH_SIZE and all of them on WRITE_SIZE
    _csv(fetch, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 1024.0 * 1024 / 2),
                               ("cwa::k_density(a, b)", 100.0), ("cwa::k_density(a, b)", 300.0)])
    _csv(write, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1024.0 * 1024),
                               ("cwa::k_density(a, b)", 50.0), ("cwa::k_density(a, b)", 50.0)])

    class A:
        write_traffic = (str(fetch), str(write))
        copy_mib = 1024

    assert b.write_traffic(A) == 0
    d = json.loads((tmp_path / "profiles" / "traffic.json").read_text())
    assert d["_calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert d["_calibration"]["write_factor"] == pytest.approx(1.0)
    dens = d["dam_break_1048576_2048"]["density"]
    assert dens["bytes"] == pytest.approx((200.0 * 2.0 + 50.0) * 1024.0)
    assert dens["dispatches"] == 2
    val, note = b.stored_traffic("dam_break_1048576_2048", "density")
    assert val == pytest.approx(dens["bytes"]) and "these kernel sources" in note
    src.write_text("// kernel v2\n")
    val, note = b.stored_traffic("dam_break_1048576_2048", "density")
    assert val is None and "other kernel sources" in note


def test_search_stats_against_brute_force(oracle):
    """On the 4 096-particle dam break: every particle's candidates include
    all its neighbours, so the mean candidate count is at least the mean
    neighbour count of brute-force all pairs; the per-wave loop (the longest
    range of each column over 64 sorted particles) is at least the mean."""
    from coupledwateranimation_amd import configs

    b = _bench()
    spec = configs.dam_break_small()
    nb = oracle.neighbor_counts(spec.params, spec.pos, all_pairs=True)
    s = b.search_stats(spec.params, spec.pos, nb)
    assert s["candidates_per_particle"] >= nb.mean()
    assert s["wave_loop_per_64"] >= s["candidates_per_particle"]
    assert s["neighbours_per_particle"]["max"] == int(nb.max())
    # a lone particle has exactly itself as candidate
    one = np.array([[0.5, 0.5, 0.5, 0.0]], np.float32)
    s1 = b.search_stats(spec.params, one, np.array([1]))
    assert s1["candidates_per_particle"] == 1.0


def test_fixed_batched_tuning_parsing():
    bench = _bench()
    tuned = {"block_density": 512, "block_force": 512, "grid_density": 1024,
             "grid_force": 1024, "xcd_queues": 1, "wave_block": 1024, "ms_per_step": 2.47}
    d = bench.fixed_tuning(json.dumps(tuned))
    assert "ms_per_step" not in d and d["grid_force"] == 1024 and len(d) == 6
    assert bench.fixed_tuning('{"xcd_queues": 0}') == {"xcd_queues": 0}
    for bad in ('[512]', '{"block": 512}', '{"grid_force": 1.5}'):
        with pytest.raises(ValueError):
            bench.fixed_tuning(bad)
