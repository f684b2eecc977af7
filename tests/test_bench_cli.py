"""bench.py's command-line contract on a machine without a GPU: more than
one GPU is refused with a message (this benchmark measures one device), and
without a device it exits non-zero instead of printing a result line."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")   # no device, also on a GPU machine
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=600, env=env, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_no_device_no_result_line():
    r = _run("--steps", "1", "--warmup", "0", "--no-extras", "--no-cpu")
    assert r.returncode != 0
    assert "no GPU" in r.stderr
    assert r.stdout.strip() == ""
