"""Build tests/cpp/replay_main (host C++ linked against libpfpa.so).  Run by
__graft_entry__.build() after the library is built."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIBDIR = os.path.join(ROOT, "phasefield-precipitate-aging_amd", "pfpa")
OUT = os.path.join(HERE, "replay_main")


def build():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    cmd = [hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "replay_main.cpp"), "-L" + LIBDIR, "-lpfpa",
           "-Wl,-rpath,$ORIGIN/../../phasefield-precipitate-aging_amd/pfpa", "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("harness build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
