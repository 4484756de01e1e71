"""Builds the oracle (CPU restatement, test infrastructure) into oracle/_build/."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")
SRC = ["tt_oracle.c", "tt_oracle_render.c"]
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-Wall", "-D_GNU_SOURCE"]


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SRC if os.path.exists(os.path.join(HERE, s))]
    tmp = OUT + ".tmp"
    r = subprocess.run(["gcc"] + CFLAGS + ["-shared", "-o", tmp] + srcs + ["-lpthread", "-lm"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
