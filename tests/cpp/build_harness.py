"""Build the C++ drivers of the drop-in API under tests/cpp (replay_main,
grouping_main; host C++ linked against libpfpa.so).  Run by
__graft_entry__.build() after the library is built."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIBDIR = os.path.join(ROOT, "phasefield-precipitate-aging_amd", "pfpa")


def _build_one(name):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = os.path.join(HERE, name)
    cmd = [hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, name + ".cpp"), "-L" + LIBDIR, "-lpfpa",
           "-Wl,-rpath,$ORIGIN/../../phasefield-precipitate-aging_amd/pfpa", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("harness build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(out + ".tmp", out)
    return out


def build():
    _build_one("grouping_main")
    return _build_one("replay_main")


if __name__ == "__main__":
    print(build())
