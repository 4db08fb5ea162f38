"""Writes flop_counts.json: exact floating-point operation counts of the DG
right-hand sides per element of a periodic brick, from the oracle run with a
counting scalar type (oracle/flop_count.cpp).  These are the algorithmic flops
of the roofline for the benchmark workloads and the other test meshes."""
import json
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[2]
EXE = ROOT / "oracle" / "_build" / "flop_count"
CASES = [("gh", (8, 8, 8)), ("gh", (6, 6, 6)), ("gh", (5, 5, 5)), ("gh", (4, 6, 5)),
         ("sw", (5, 5, 5)), ("sw", (6, 6, 6)), ("sw", (8, 8, 8))]


def count(system, extents):
    out = subprocess.run([str(EXE), system, *map(str, extents)], check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out)


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    data = [count(s, e) for s, e in CASES]
    (pathlib.Path(__file__).parent / "flop_counts.json").write_text(json.dumps(data, indent=1) + "\n")


if __name__ == "__main__":
    main()
