"""Writes the golden .vibebin files with the reference's own BinaryWriter.

oracle/ref.mk compiles the reference's src/io/binary_writer.cpp and
src/grid/geometry.cpp (unmodified, from where they lie under /root/reference)
with oracle/ref_vibebin.cpp into oracle/_ref/ref_vibebin; this script runs it
and stores the three files it writes here (a few KB each; the content is
described in oracle/ref_vibebin.cpp).
Run:  make -C oracle -f ref.mk && python tests/golden/make_vibebin_golden.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "ref_vibebin")


def main():
    assert os.path.exists(EXE), "build oracle/_ref first: make -C oracle -f ref.mk"
    for mode in ("generic", "restart", "restart13"):
        out = os.path.join(HERE, f"vibebin_{mode}.vibebin")
        subprocess.run([EXE, mode, out], check=True)
        print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
