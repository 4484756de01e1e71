"""Builds the oracle (CPU restatement, test infrastructure) into oracle/_build/."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")
SRC = ["tt_oracle.c", "tt_oracle_render.c"]
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-Wall", "-D_GNU_SOURCE"]


def build(plant=0):
    """Builds liboracle.so; plant = k > 0 builds liboracle_plant<k>.so with the deliberate
    mistake k compiled in (tests only: their checks must fail on it)."""
    out = OUT if not plant else os.path.join(HERE, "_build", f"liboracle_plant{int(plant)}.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SRC if os.path.exists(os.path.join(HERE, s))]
    tmp = out + f".tmp{os.getpid()}"
    flags = CFLAGS + ([f"-DO_PLANT={int(plant)}"] if plant else [])
    r = subprocess.run(["gcc"] + flags + ["-shared", "-o", tmp] + srcs + ["-lpthread", "-lm"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
