"""Compile the C restatements (fe_oracle.c, mr_oracle.c) into oracle/liboracle_fe.so.

Test infrastructure: used only by tests/ and bench.py's cpu_baseline leg.
Built for x86-64-v3 so the library runs on any AVX2 host it is shipped to.
"""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle_fe.so"


def build():
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC", str(HERE / "fe_oracle.c"),
           str(HERE / "mr_oracle.c"), "-o", str(SO), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
    return SO


if __name__ == "__main__":
    try:
        build()
        print(f"built {SO.name}")
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
