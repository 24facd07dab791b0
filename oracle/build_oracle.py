"""Builds the oracle's C restatement (simp_cpu.c) into oracle/_build/.

TEST / BASELINE INFRASTRUCTURE ONLY. gcc, -O3, OpenMP, x86-64-v3 (AVX2+FMA)
so the object runs on any recent x86 host.
"""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "libsimp_cpu.so"


def build() -> Path:
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC", "-std=c11",
           str(HERE / "simp_cpu.c"), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build())
