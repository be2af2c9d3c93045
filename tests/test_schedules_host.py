"""The autotuner's schedule table on the host (no device): JSON round trip, and rejection of
malformed files and of schedules this build does not offer."""
import json

import pytest


def _entry(**kw):
    e = {"model": 0, "nlayers": 1, "nx": 256, "ny": 256, "dtype": 0, "p": [4, 8, 4], "step_us": 81.25}
    e.update(kw)
    return e


def test_round_trip(gf, tmp_path):
    gf._lib.lib.gf_schedules_clear()
    src = tmp_path / "in.json"
    src.write_text(json.dumps({"format": "gf-schedules-1",
                               "entries": [_entry(split=[1, 0]), _entry(nx=64, ny=32, p=[4, 8, 0], dtype=1),
                                           _entry(nx=512, ny=512, split=[1, 1, 1]),
                                           _entry(nx=1024, ny=1024, p=[8, 8, 8], split=[1, 2, 1])]}))
    assert gf.load_schedules(src) == 4
    out = tmp_path / "out.json"
    gf.save_schedules(out)
    got = json.loads(out.read_text())
    assert got["format"] == "gf-schedules-1"
    ents = sorted(got["entries"], key=lambda e: e["nx"])
    assert [e["p"] for e in ents] == [[4, 8, 0], [4, 8, 4], [4, 8, 4], [8, 8, 8]]
    assert ents[1]["step_us"] == 81.25 and ents[0]["ny"] == 32 and ents[0]["dtype"] == 1
    # absent: no splits; two entries: K3's column split off
    assert ents[1]["split"] == [1, 0, 0] and ents[0]["split"] == [0, 0, 0] and ents[2]["split"] == [1, 1, 1]
    assert ents[3]["split"] == [1, 2, 1]     # K2's staged pass
    gf._lib.lib.gf_schedules_clear()
    gf.save_schedules(out)
    assert json.loads(out.read_text())["entries"] == []


@pytest.mark.parametrize("text,msg", [
    ("{", "expected a string"),
    ("[1, 2", "expected .]."),
    ('{"format": "other", "entries": []}', "not a gf-schedules-1 file"),
    ('{"format": "gf-schedules-1", "entries": {}}', "must be an array"),
    ('{"format": "gf-schedules-1", "entries": [{"model": 0}]}', "must be an integer"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(p=[16, 8, 8])]}), "does not offer"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(nx=16, p=[4, 8, 4])]}), "does not offer"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(p=[4, 8])]}), "3 integers"),
    ('{"format": "gf-schedules-1", "entries": []} x', "trailing"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(split=[2, 0])]}), "0 or 1"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(split=[0, 0, 3])]}), "0 or 1"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(split=[0, 3, 0])]}), "0, 1 or 2 for K2"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(split=[0, 0, 2])]}), "0 or 1"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(split=[0, 0, 0, 1])]}), "2 or 3 integers"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(nlayers=2, p=[0, 0, 0], split=[0, 0, 1])]}), "single-layer"),
    (json.dumps({"format": "gf-schedules-1", "entries": [_entry(nlayers=3, p=[0, 0, 0], split=[1, 0])]}), "single-layer"),
])
def test_malformed_tables_are_rejected(gf, tmp_path, text, msg):
    gf._lib.lib.gf_schedules_clear()
    f = tmp_path / "bad.json"
    f.write_text(text)
    with pytest.raises(gf._lib.GFError, match=msg):
        gf.load_schedules(f)
    out = tmp_path / "after.json"
    gf.save_schedules(out)
    assert json.loads(out.read_text())["entries"] == []     # nothing of a bad file is kept


def test_missing_file(gf, tmp_path):
    with pytest.raises(gf._lib.GFError, match="cannot open"):
        gf.load_schedules(tmp_path / "none.json")


def test_shipped_table_loads_with_every_entry_offered(gf, tmp_path):
    # the table beside the library is read at start-up; a malformed or unoffered entry would
    # leave every problem on the built-in schedules without an error
    import os
    path = os.path.join(os.path.dirname(gf.__file__), "schedules.json")
    ents = json.load(open(path))["entries"]
    gf._lib.lib.gf_schedules_clear()
    assert gf.load_schedules(path) == len(ents) > 0
    head = [e for e in ents if (e["model"], e["nlayers"], e["nx"], e["ny"], e["dtype"]) == (0, 1, 2048, 2048, 0)]
    assert len(head) == 1 and len(head[0]["split"]) == 3      # the headline workload is tuned
    out = tmp_path / "again.json"
    gf.save_schedules(out)
    assert sorted(json.dumps(e, sort_keys=True) for e in json.loads(out.read_text())["entries"]) == \
        sorted(json.dumps(e, sort_keys=True) for e in ents)
    gf._lib.lib.gf_schedules_clear()
