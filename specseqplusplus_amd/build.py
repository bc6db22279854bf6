"""Build libssq.so and the specseq++ CLI in-tree (hipcc, gfx950)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8):
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", os.path.join(HERE, "csrc"), "all"])
    return os.path.join(HERE, "lib", "libssq.so")


if __name__ == "__main__":
    print(build())
