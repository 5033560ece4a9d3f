"""bench.py contract on the CPU: it refuses more than one GPU with a message,
and its algorithmic byte accounting is the per-trial model (first FFT pass
reads the series through the resampler and writes n/2 complex; every other
pass reads and writes n/2 complex; the spectrum pass reads n/2 complex and
writes n/2+1 floats; the harmonic pass reads n/2+1 floats)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_stage_bytes_model():
    sys.path.insert(0, ROOT)
    import bench
    n = 1 << 23
    m, nb = n // 2, n // 2 + 1
    two = bench.stage_bytes(n, 2, bound=False)
    assert two == [4 * n + 8 * m, 16 * m, 8 * m + 4 * nb, 4 * nb]
    three = bench.stage_bytes(n, 3, bound=False)
    assert three[:3] == [4 * n + 8 * m, 16 * m, 16 * m] and three[3:] == two[2:]
    # the fused trial moves 24n bytes with two FFT passes (about 28n with three)
    assert abs(sum(two) - 24 * n) <= 16
    # bound-filtered harmonic form: one 16-bit bound per two bins (n/2 + 2
    # bytes), read once by the harmonic pass; after the natural layout a
    # pooling pass reads the spectrum again to write them
    qb = nb + 1
    bound = bench.stage_bytes(n, 2, fused=False)
    assert bound[:2] == two[:2]
    assert bound[2] == 8 * m + 8 * nb + qb and bound[3] == qb
    # fused into the row-layout spectrum pass, which then writes no spectrum
    fused = bench.stage_bytes(n, 2)
    assert fused[2] == 8 * m + qb and fused[3] == qb


def test_cpu_baseline_threads_and_model():
    """The CPU baseline uses every CPU this process may run on, up to the
    16-CPU share of a one-GPU host, and names the CPU model."""
    sys.path.insert(0, ROOT)
    import bench
    n = bench.cpu_threads()
    assert 1 <= n <= 16
    assert n == min(16, len(os.sched_getaffinity(0)))
    assert isinstance(bench.cpu_model(), str) and bench.cpu_model()


def test_serving_stream_shape():
    """The default serving stream is the serving workload's: 64 requests of
    16..256 DM trials."""
    sys.path.insert(0, ROOT)
    import inspect
    import bench
    sig = inspect.signature(bench.serve_stream)
    assert sig.parameters["nreq"].default == 64
    assert sig.parameters["dms"].default == (16, 256)
    assert sig.parameters["rate"].default == 4.0
