"""The cylinder batch table: the case layer's MT19937-64 against the C++ standard library.

batch_cases.csv was written by tests/golden/gen_batch_cases.cpp with std::mt19937_64, so the
table the batch benchmark builds in Python is pinned to the standard engine, and the engine
itself is pinned by the standard's known answer (the 10000th output of the default seed).
"""
import csv
import os

from hyperboliconvergence_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt19937_64_known_answer():
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed mt19937_64
    r = cases.MT19937_64()
    for _ in range(9999):
        r()
    assert r() == 9981545732273789042


def test_batch_table_matches_std_engine_fixture():
    with open(os.path.join(HERE, "golden", "batch_cases.csv")) as f:
        rows = list(csv.DictReader(f))
    table = cases.batch_case_table()
    assert len(rows) == len(table) == 96
    for r, t in zip(rows, table):
        assert int(r["case"]) == t["case"]
        assert int(r["size"]) == t["size"]
        assert (int(r["nth"]), int(r["nr"])) == (t["nth"], t["nr"])
        assert float(r["mach"]) == t["mach"]
        assert float(r["alpha_deg"]) == t["alpha_deg"]
    assert all(0.2 <= t["mach"] < 0.8 and -5.0 <= t["alpha_deg"] < 5.0 for t in table)


def test_batch_cell_count():
    table = cases.batch_case_table()
    assert sum(t["nth"] * t["nr"] for t in table) == 5496832
