"""Builds the oracle's C restatements (test infrastructure) into oracle/_c/ with gcc."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build():
    out = HERE / "_c"
    out.mkdir(exist_ok=True)
    built = []
    for src in sorted(HERE.glob("*.c")):
        so = out / (src.stem + ".so")
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-o", str(so), str(src), "-lm"]
        subprocess.run(cmd, check=True)
        built.append(so)
    return built


if __name__ == "__main__":
    print(build())
