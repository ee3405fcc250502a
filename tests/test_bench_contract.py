"""CPU checks of bench.py's host logic: the counter-summary lookup refuses a
summary taken from other kernel sources, and multi-GPU runs are refused."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_lookup_needs_the_current_kernel_sources(tmp_path, monkeypatch):
    f = tmp_path / 'traffic.txt'
    digest = bench.kernel_source_digest()
    assert len(digest) == 16
    f.write_text('source_digest = %s\ntraffic[k<double>] = 1234\n' % digest)
    monkeypatch.setattr(bench, 'TRAFFIC_FILE', str(f))
    assert bench.measured_traffic('k<double>') == 1234
    assert bench.measured_traffic('other<double>') is None
    f.write_text('source_digest = 0000000000000000\ntraffic[k<double>] = 1234\n')
    assert bench.measured_traffic('k<double>') is None      # stale summary
    f.write_text('traffic[k<double>] = 1234\n')
    assert bench.measured_traffic('k<double>') is None      # no digest at all


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and 'one GPU' in r.stdout
