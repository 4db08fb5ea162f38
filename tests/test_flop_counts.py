"""Algorithmic flop counts of the right-hand sides (the oracle run with a
counting scalar type, oracle/flop_count.cpp) match the committed fixture, and
the counts behave as the operation structure says they must."""
import json
import pathlib
import subprocess

import oracle_lib as ol

GOLDEN = pathlib.Path(__file__).parent / "golden" / "flop_counts.json"
EXE = ol.ROOT / "oracle" / "_build" / "flop_count"


def count(system, extents):
    if not EXE.exists():
        ol.build()
    out = subprocess.run([str(EXE), system, *map(str, extents)], check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out)


def test_counts_match_fixture():
    fixture = json.loads(GOLDEN.read_text())
    for entry in fixture:
        if entry["system"] == "sw" or entry["extents"] == [5, 5, 5]:
            assert count(entry["system"], entry["extents"]) == entry


def test_one_sqrt_per_face_point_side():
    # ScalarWave: every face point normalises its own and the neighbour's
    # normal (one sqrt each); the volume terms take no square roots.
    for n in (3, 5, 6):
        c = count("sw", (n, n, n))
        assert c["sqrt"] == 2 * 6 * n * n
