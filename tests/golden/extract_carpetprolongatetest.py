"""Extracts the CarpetProlongateTest golden outputs used by the oracle pins.

Reads the reference's stored test outputs (data files under
CarpetProlongateTest/test/*/) and writes tests/golden/carpetprolongatetest.json:
  - errornorm (max |u - exact| over all levels) of every test that has one;
  - every refined-level sample of the 1-D x/y/z/diagonal lines whose
    difference du = u - exact is non-zero: (test, line, x, y, z, u0, du).
Run in a container that has the reference tree at /root/reference.
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "CarpetProlongateTest" / "test"
OUT = Path(__file__).resolve().parent / "carpetprolongatetest.json"


def rows(path):
    for line in path.read_text().splitlines():
        if not line.strip() or line.startswith("#"):
            continue
        yield line.split()


def main():
    res = {"source": "CarpetProlongateTest/test/<test>/carpetprolongatetest-*.asc", "tests": {}}
    for d in sorted(p for p in REF.iterdir() if p.is_dir()):
        t = {}
        en = d / "carpetprolongatetest-errornorm..asc"
        if en.exists():
            vals = [float(r[-1]) for r in rows(en)]
            t["errornorm"] = vals[-1]
        samples = []
        for line in ("x", "y", "z", "d"):
            f = d / f"carpetprolongatetest-difference.{line}.asc"
            if not f.exists():
                continue
            for r in rows(f):
                # it tl rl c ml ix iy iz time x y z u0 du
                rl = int(r[2])
                du = float(r[13])
                if rl >= 1 and du != 0.0:
                    samples.append({"line": line, "rl": rl, "c": int(r[3]),
                                    "x": float(r[9]), "y": float(r[10]), "z": float(r[11]),
                                    "u0": float(r[12]), "du": du})
        t["samples"] = samples
        if en.exists() or samples:
            res["tests"][d.name] = t
    OUT.write_text(json.dumps(res, indent=1, sort_keys=True))
    print(f"wrote {OUT}: {len(res['tests'])} tests")


if __name__ == "__main__":
    main()
