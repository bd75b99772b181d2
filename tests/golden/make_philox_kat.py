"""Writes tests/golden/philox_kat.json.

Two independent sources of Philox4x32-10 known answers:
  * the Random123 published known-answer vectors (Salmon et al., SC'11,
    distributed with the Random123 library as kat_vectors), typed in below;
  * rocRAND's open-source engine, compiled on the host from
    philox_rocrand_kat.cpp with hipcc.
Run: python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

RANDOM123 = [
    {"key": [0x00000000, 0x00000000], "ctr": [0, 0, 0, 0],
     "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
    {"key": [0xFFFFFFFF, 0xFFFFFFFF], "ctr": [0xFFFFFFFF] * 4,
     "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
    {"key": [0xA4093822, 0x299F31D0], "ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
     "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
]


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "kat")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", os.path.join(HERE, "philox_rocrand_kat.cpp"),
                        "-o", exe], check=True)
        roc = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    out = {"random123": RANDOM123, "rocrand": roc}
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(RANDOM123)} Random123 and {len(roc)} rocRAND vectors")


if __name__ == "__main__":
    main()
