"""Copies the reference's golden fixtures (tests/refs/*.npz of zoomy-amrex,
data files: final Q and Qaux of its six end-to-end cases) into
tests/golden/refs/ so that tests never read the reference tree at run time.

Run from the repository root:  python tests/golden/copy_reference_fixtures.py
"""
import pathlib
import shutil

SRC = pathlib.Path("/root/reference/tests/refs")
DST = pathlib.Path(__file__).resolve().parent / "refs"

if __name__ == "__main__":
    DST.mkdir(exist_ok=True)
    for f in sorted(SRC.glob("*.npz")):
        shutil.copy2(f, DST / f.name)
        print("copied", f.name)
