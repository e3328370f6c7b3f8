"""bench.py's command line on the CPU: a request for several GPUs is refused
with a message, and without a GPU the benchmark fails loudly in the native
library instead of timing a CPU stand-in."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def test_several_gpus_are_refused():
    res = _bench("--gpus", "2")
    assert res.returncode == 2
    assert "one GPU only" in res.stdout + res.stderr


def test_no_gpu_fails_in_the_native_library():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = os.path.join(ROOT, "axi_visualizer_amd", "libaxr.so")
    if not os.path.exists(lib):
        pytest.skip("libaxr.so is not built")
    res = _bench("--config", "c1", "--steps", "1", "--warmup", "0", "--no-cpu-baseline")
    assert res.returncode != 0
    assert "AxrError" in res.stderr and "axr_create failed" in res.stderr
    assert '"metric"' not in res.stdout  # no result line from a stand-in
