import fcntl
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    # the oracle is the checker of every test: a plain gcc build of one C file
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        make_locked("-s", "-C", os.path.join(ROOT, "oracle"))


def make_locked(*args):
    """`make *args` under a file lock, so that test workers running in
    parallel never load a library another worker is relinking."""
    os.makedirs(os.path.join(ROOT, "biospring_amd", "csrc", "_build"), exist_ok=True)
    with open(os.path.join(ROOT, "biospring_amd", "csrc", "_build", ".make.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        try:
            subprocess.run(["make", *args], check=True, stdout=subprocess.DEVNULL)
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)
