"""bench.py's command line without a GPU: a run on more than one GPU is refused
with a message (the engine runs on one), before anything touches a device."""
import os
import subprocess
import sys

from conftest import REPO


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
