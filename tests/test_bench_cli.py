"""bench.py's command line without a GPU: one GPU only, and the serving
step's request accounting."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_default_run_is_the_serving_trace_of_4000_requests():
    sys.path.insert(0, ROOT)
    import bench
    args = bench.make_parser().parse_args([])
    assert args.gpus == 1 and args.workload == "serving" and args.requests == 0
    # a serving step is REQUESTS_PER_STEP requests: 4000 timed, 200 warm-up by default
    assert bench.REQUESTS_PER_STEP * args.steps == 4000
    assert bench.REQUESTS_PER_STEP * args.warmup == 200
