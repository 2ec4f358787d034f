"""CPU: bench.py's command-line contract that needs no GPU -- it measures one
GPU only and says so (exit status 2) for --gpus above 1 or a multi-rank
launch, before it touches the GPU."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py")] + args, capture_output=True, text=True,
                          env=env, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run_bench(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_multi_rank_launch_is_refused():
    r = run_bench(["--gpus", "1"], {"WORLD_SIZE": "2", "RANK": "0"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_help_lists_the_sections():
    r = run_bench(["--help"])
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--no-serving", "--serve-offered-10m", "--serve-offered-100m"):
        assert flag in r.stdout
