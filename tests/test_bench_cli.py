"""bench.py's command-line contract, without a device: more than one GPU is
refused with a message before anything touches the GPU, and every option the
measurement documents is accepted by the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_documented_options_parse():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--config", "--batch-arms", "--solid-size", "--plane-size", "--load-hint",
                "--ru-delta", "--profile-spmv", "--profile-shape", "--serve-precision", "--serve-oracle-dof"):
        assert opt in r.stdout, opt
    for arm in ("fp32_ir_batched_mf", "fp32_ir_batched", "fp64_batched", "fp64_batched_mf", "fp64_sequential"):
        assert arm in r.stdout, arm
