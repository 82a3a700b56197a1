"""The autotuner's output: the table compiled into the library is the one its
committed measurement record chose, the lookup takes the row of the largest
minimum atom count not above a system's, and the generated header round-trips
(no GPU: the tuner's measurements ran on the MI355X, tuning.json keeps them)."""
import json
import os

import pytest

import rainmd_amd as R
from rainmd_amd import autotune

HERE = os.path.dirname(os.path.abspath(__file__))


def _record():
    with open(autotune.TABLE_JSON) as f:
        return json.load(f)


def test_compiled_table_is_the_record_s_choice():
    record = _record()
    assert record["classes"], "no tuned classes"
    for c in record["classes"]:
        kind = "single" if c["kind"] == 0 else "batch"
        shape = R.tuned_shape(kind, c["atoms"] or 0)
        best = c["best"]
        print(c["label"], shape, best)
        assert shape["lanes_per_atom"] == best["lanes"]
        assert shape["cell_reach"] == best["cell_reach"]
        assert shape["search_mode"] == best["search_mode"]
        assert shape["skin"] == pytest.approx(best["skin"])
        if kind == "single":
            assert shape["pme_bins"] == best["pme_bins"]
        # the choice is the fastest trial the record holds for the class
        fastest = min(c["trials"], key=lambda t: t["ms_per_step"])
        assert fastest["params"] == best and fastest["ms_per_step"] == c["best_ms_per_step"]


def test_rows_are_looked_up_by_atom_count():
    rows = [r for r in autotune.current_table() if r["kind"] == 0]
    starts = sorted(r["minAtoms"] for r in rows)
    for lo, hi in zip(starts, starts[1:] + [10 ** 7]):
        row = next(r for r in rows if r["minAtoms"] == lo)
        for atoms in (lo, (lo + hi) // 2, hi - 1):
            assert R.tuned_shape("single", atoms)["lanes_per_atom"] == row["lanes"], atoms


def test_header_round_trips():
    rows = autotune.current_table()
    with open(autotune.TABLE_HEADER) as f:
        assert f.read() == autotune.header(rows)
