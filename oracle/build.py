"""Build the oracle's C restatements into ``oracle/_build/liboracle.so`` (make)."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "_build" / "liboracle.so"


def build():
    subprocess.run(["make", "-s", "-C", str(HERE), "-j4"], check=True)
    return LIBRARY


if __name__ == "__main__":
    print(build())
