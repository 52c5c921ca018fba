"""Builds the native parts in-tree: planc_amd/libplanc_hip.so (hipcc, gfx950)
and, when present, the C++ host library / CLI of planc_amd/host."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def _make(path, jobs=8):
    subprocess.run(["make", "-C", path, f"-j{jobs}"], check=True)


def build(jobs=8):
    _make(os.path.join(HERE, "csrc"), jobs)
    host = os.path.join(HERE, "host")
    if os.path.exists(os.path.join(host, "Makefile")):
        _make(host, jobs)


if __name__ == "__main__":
    build()
