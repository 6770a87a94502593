"""The batch-plan tuning table (include/lsdg_tune.h) on the CPU: entries
round-trip through the JSON file, a file merges into the table, malformed
files and out-of-range shapes are refused, and a file written by hand in the
documented format loads."""
import json

import pytest

import lsdg

SHAPE = dict(schedule=lsdg.SCHEDULE_COOP, coop_block=256, coop_k=8, solo_max=4096, job_max_points=6144, chains=1,
             points_per_thread=4, xcd_groups=8)


@pytest.fixture(autouse=True)
def clean_table():
    lsdg.tune_table_clear()
    yield
    lsdg.tune_table_clear()


def test_round_trip(tmp_path):
    steps = dict(SHAPE, schedule=lsdg.SCHEDULE_STEPS, chains=3, points_per_thread=8, xcd_groups=1)
    lsdg.tune_table_put("se3/fp32/p19/single/n6", SHAPE)
    lsdg.tune_table_put("sim3/fp16/p21/single/n7", steps)
    path = tmp_path / "tune.json"
    lsdg.tune_table_save(path)
    doc = json.loads(path.read_text())
    assert doc["lsdg_tune"] == 1 and len(doc["entries"]) == 2
    lsdg.tune_table_clear()
    assert lsdg.tune_table_size() == 0 and lsdg.tune_table_get("se3/fp32/p19/single/n6") is None
    lsdg.tune_table_load(path)
    assert lsdg.tune_table_get("se3/fp32/p19/single/n6") == SHAPE
    assert lsdg.tune_table_get("sim3/fp16/p21/single/n7") == steps


def test_load_merges_and_file_wins(tmp_path):
    lsdg.tune_table_put("a/b", SHAPE)
    lsdg.tune_table_put("keep/me", SHAPE)
    other = dict(SHAPE, coop_k=2)
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"lsdg_tune": 1, "comment": ["any", {"extra": 1}],
                                "entries": [dict(key="a/b", **other)]}))
    lsdg.tune_table_load(path)
    assert lsdg.tune_table_get("a/b") == other and lsdg.tune_table_get("keep/me") == SHAPE


@pytest.mark.parametrize("text", ['{"entries": [{"key": "x", "schedule": 0', '{"entries": [1, 2,]}', "not json"])
def test_malformed_files_are_refused(tmp_path, text):
    path = tmp_path / "bad.json"
    path.write_text(text)
    with pytest.raises(lsdg.LsdgError):
        lsdg.tune_table_load(path)
    assert lsdg.tune_table_size() == 0


def test_bad_shapes_and_keys_are_refused(tmp_path):
    for bad in (dict(SHAPE, coop_block=384), dict(SHAPE, chains=0), dict(SHAPE, schedule=5),
                dict(SHAPE, xcd_groups=0)):
        with pytest.raises(lsdg.LsdgError):
            lsdg.tune_table_put("k", bad)
    for key in ("", "has space", "x" * 64, 'quo"te'):
        with pytest.raises(lsdg.LsdgError):
            lsdg.tune_table_put(key, SHAPE)
    path = tmp_path / "b.json"
    path.write_text(json.dumps({"entries": [dict(key="ok", **dict(SHAPE, coop_block=100))]}))
    with pytest.raises(lsdg.LsdgError):
        lsdg.tune_table_load(path)
    assert lsdg.tune_table_size() == 0
