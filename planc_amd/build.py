"""Builds the native parts in-tree: planc_amd/libplanc_hip.so (hipcc, gfx950)
and, when present, the C++ host library / CLI of planc_amd/host."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def _make(path, jobs=8, force=False):
    cmd = ["make", "-C", path, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


def build(jobs=8, force=False):
    """force=True recompiles every source even where the outputs are newer."""
    _make(os.path.join(HERE, "csrc"), jobs, force)
    host = os.path.join(HERE, "host")
    if os.path.exists(os.path.join(host, "Makefile")):
        _make(host, jobs, force)


if __name__ == "__main__":
    build()
