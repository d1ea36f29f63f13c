"""bench.py's contract pieces that need no GPU: one GPU only (--gpus N > 1
exits with a message before touching a device), and the roofline's traffic
figure comes from the committed rocprofv3 summary, consistent with the
counter rows it lists."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_traffic_summary_parses_and_matches_its_counters():
    sys.path.insert(0, ROOT)
    import bench

    traffic, src = bench.measured_traffic()
    assert src == os.path.join("profiles", "k_backward_batched_std_traffic.md")
    text = open(os.path.join(ROOT, src)).read()

    def avg(kernel, counter):
        m = re.search(r"\| %s \| \d+ \| %s \| \d+ \| ([0-9.]+) \|" % (kernel, counter), text)
        return float(m.group(1))

    kb, copy = r"k_backward<4, false, \d+>", "__amd_rocclr_copyBuffer"
    # bytes = FETCH_SIZE x f + WRITE_SIZE x w (KiB), with f and w the factors
    # that make the 256 MiB copy's counters read 262144 KiB each way
    f = 262144.0 / avg(copy, "FETCH_SIZE")
    w = 262144.0 / avg(copy, "WRITE_SIZE")
    expect = (avg(kb, "FETCH_SIZE") * f + avg(kb, "WRITE_SIZE") * w) * 1024
    assert abs(traffic - expect) <= 1.0
    # at least the algorithmic bytes of the launch: (9 + 2/5) images x P K N x 4 B
    assert traffic >= (9 + 2 / 5) * 64 * 5 * 16384 * 4
