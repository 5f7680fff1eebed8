"""CPU checks of the tunables registry and its JSON cache (no GPU calls)."""
import json

import pytest


@pytest.fixture(scope="module")
def at():
    import frovedis_amd
    frovedis_amd.lib()
    from frovedis_amd import autotune
    return autotune


def test_set_get_default(at):
    at.tune_set("test.a", None)
    assert at.tune_get("test.a", 17) == 17
    at.tune_set("test.a", 5)
    assert at.tune_get("test.a", 17) == 5
    at.tune_set("test.a", None)
    assert at.tune_get("test.a", 17) == 17


def test_save_load_roundtrip(at, tmp_path):
    at.tune_set("groupby.slots", 1024)
    at.tune_set("groupby.fill_pct", 40)
    p = tmp_path / "tune.json"
    at.save(p)
    saved = json.loads(p.read_text())
    assert saved["groupby.slots"] == 1024 and saved["groupby.fill_pct"] == 40
    at.tune_set("groupby.slots", None)
    at.tune_set("groupby.fill_pct", None)
    assert at.tune_get("groupby.slots", -7) == -7
    assert at.load(p)
    assert at.tune_get("groupby.slots") == 1024 and at.tune_get("groupby.fill_pct") == 40
    at.tune_set("groupby.slots", None)
    at.tune_set("groupby.fill_pct", None)
    assert not at.load(tmp_path / "missing.json")


def test_load_rejects_malformed(at, tmp_path):
    import frovedis_amd
    p = tmp_path / "bad.json"
    p.write_text('{"groupby.slots": "many"}')
    with pytest.raises(frovedis_amd.FhipError):
        at.load(p)


def test_batch_size_class(at):
    assert at.batch_size_class(4097) == 13
    assert at.batch_size_class(8192) == 13
    assert at.batch_size_class(8193) == 14
    assert at.batch_size_class(1 << 20) == 20
