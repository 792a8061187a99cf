"""bench.py's command-line contract, checked without a GPU: more than one
GPU is refused with a message, and without a GPU the benchmark fails
instead of printing a number from some other path."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(*args, timeout=600):
    return subprocess.run([sys.executable, BENCH, *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stdout + r.stderr
    assert "{" not in r.stdout


def _gpu_visible():
    import torch
    return torch.cuda.device_count() > 0


@pytest.mark.skipif(_gpu_visible(), reason="checks the behaviour on a host without a GPU")
def test_no_gpu_fails_without_a_result():
    r = _run("--steps", "1", "--warmup", "0", "--cpu-steps", "0")
    assert r.returncode != 0
    assert "{" not in r.stdout


def test_ensemble_cpu_baseline_runs_uneven_members_on_the_oracle():
    """The ensemble's CPU leg runs members whose planes are not square
    (np > mp + 1 among them) through the C oracle; the reference build is
    kept to square planes, where its y-wall zeroing stays in bounds."""
    sys.path.insert(0, ROOT)
    import bench
    from advectionhpctester_amd import testdata
    shapes = [(48, 16, 8), (16, 32, 8), (24, 24, 8)]
    mem = [testdata.ensemble_member(*sh, seed=1000 + q) for q, sh in enumerate(shapes)]
    x0 = [m["x"] for m in mem]
    rest = [[m[k] for m in mem] for k in ("u1", "u2", "u3", "rhr", "h", "bcx", "bcy")]
    out = bench.ensemble_cpu_baseline(x0, rest, shapes, max_cells=10 ** 9)
    assert out["kind"] == "port" and out["cores"] == 1 and out["value"] > 0
    assert "3 of the 3 members" in out["sample"]
