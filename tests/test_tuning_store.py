"""The launch-shape store (slimtracin_amd.tuning_store): file round trip,
layouts it refuses, keys and shape records; no GPU needed."""
from __future__ import annotations

import json
import os

import pytest

import slimtracin_amd as st
from slimtracin_amd import tuning_store as ts


def _shape(**kw):
    d = dict(tile_shift=3, blocks_per_cu=4, persistent=2, cost_order=0, ray_sort=0, stack_entries=16, primary_waves=4,
             bounce_order=1)
    d.update(kw)
    return st.Tuning(*[d[f] for f in ts.FIELDS])


def test_shape_record_round_trip():
    t = _shape(tile_shift=4, ray_sort=1)
    d = ts.tuning_to_dict(t)
    assert d == dict(tile_shift=4, blocks_per_cu=4, persistent=2, cost_order=0, ray_sort=1, stack_entries=16,
                     primary_waves=4, bounce_order=1)
    assert ts.tuning_to_dict(ts.tuning_from_dict(d)) == d


@pytest.mark.parametrize("bad", [{}, {"tile_shift": 3}, dict.fromkeys(ts.FIELDS, -1), dict.fromkeys(ts.FIELDS, 1.5),
                                 dict(dict.fromkeys(ts.FIELDS, 1), extra=1)])
def test_shape_record_rejects_other_layouts(bad):
    with pytest.raises(ValueError):
        ts.tuning_from_dict(bad)


def test_store_file_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "tuning.json")
    assert ts.load(path) == {"version": 1, "gpus": {}}
    summary = [[16, 0, 0, 0, 640, 480, 0, 2, 0]]
    wkey = ts.workload_key(summary)
    data = ts.store(ts.load(path), "gfx950/256cu/AMD Instinct MI355X", wkey, _shape(blocks_per_cu=8), 1.25, summary)
    ts.save(data, path)
    assert sorted(os.listdir(tmp_path / "sub")) == ["tuning.json"]   # no temporary file left behind
    got = ts.load(path)
    t = ts.lookup(got, "gfx950/256cu/AMD Instinct MI355X", wkey)
    assert t is not None and ts.tuning_to_dict(t) == ts.tuning_to_dict(_shape(blocks_per_cu=8))
    assert got["gpus"]["gfx950/256cu/AMD Instinct MI355X"][wkey]["ms"] == 1.25
    assert ts.lookup(got, "gfx950/128cu/other", wkey) is None
    assert ts.lookup(got, "gfx950/256cu/AMD Instinct MI355X", "0" * 20) is None


@pytest.mark.parametrize("text", ["", "not json", "[]", json.dumps({"version": 2, "gpus": {}}),
                                  json.dumps({"version": 1, "gpus": []})])
def test_unreadable_store_is_empty(tmp_path, text):
    p = tmp_path / "tuning.json"
    p.write_text(text)
    assert ts.load(str(p)) == {"version": 1, "gpus": {}}


def test_entry_with_a_bad_shape_is_not_found(tmp_path):
    data = {"version": 1, "gpus": {"g": {"w": {"tuning": {"tile_shift": 3}, "ms": 1.0}}}}
    assert ts.lookup(data, "g", "w") is None


def test_keys():
    a = [[16, 0, 0, 0, 640, 480, 0, 2, 0]]
    b = [[16, 0, 0, 0, 640, 481, 0, 2, 0]]
    assert ts.workload_key(a) == ts.workload_key([list(x) for x in a])
    assert ts.workload_key(a) != ts.workload_key(b)
    assert ts.gpu_key({"arch": "gfx950:sramecc+:xnack-", "compute_units": 256, "name": "AMD Instinct MI355X"}) == \
        "gfx950/256cu/AMD Instinct MI355X"


def test_default_path_follows_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("SLIMTRACIN_TUNING", str(tmp_path / "t.json"))
    assert ts.default_path() == str(tmp_path / "t.json")
    monkeypatch.delenv("SLIMTRACIN_TUNING")
    assert ts.default_path().endswith(os.path.join(".cache", "slimtracin_amd", "tuning.json"))
