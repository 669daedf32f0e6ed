"""Builds oracle/lib/libcpu_ref.so (gcc -O3 -fopenmp) from oracle/cpu_ref.c.

x86-64-v3 (AVX2 + FMA) is targeted so the library runs on any current x86
host; no -ffast-math, so the float64 exp and summation order stay as written.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lib", "libcpu_ref.so")


def build(verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared",
           os.path.join(HERE, "cpu_ref.c"), "-o", OUT + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"[build] wrote {OUT}", flush=True)
    return OUT


if __name__ == "__main__":
    build()
