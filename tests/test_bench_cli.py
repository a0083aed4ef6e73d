"""bench.py's command line without a GPU: more than one GPU is refused with a message (every box
here has one), and the options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_options_parse():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--workload", "--window", "--tile-autotune", "--no-cpu"):
        assert opt in r.stdout


def test_tile_autotune_default_depends_on_the_workload():
    sys.path.insert(0, ROOT)
    import bench

    assert bench.parse_args([]).tile_autotune == 0  # zone108 keeps the engine's one-row default
    assert bench.parse_args(["--workload", "zone7"]).tile_autotune == 1
    assert bench.parse_args(["--tile-autotune", "1"]).tile_autotune == 1
