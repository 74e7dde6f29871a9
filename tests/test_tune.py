"""The autotuner's host logic: candidate sets, choice and the JSON cache
(no device: a stand-in context records what the tuner sets)."""
import json

import pytest

from parafem_amd import tune as T


def test_candidates_match_the_library_limits():
    # pf_tune_set accepts powers of two up to 2 (ntot 60), 4 (ntot 24), 8 (ntot 8)
    assert T.EPL_CANDIDATES == {60: (1, 2), 24: (1, 2, 4), 8: (1, 2, 4, 8)}
    assert T.BLOCK_WAVES == (1, 2, 4)
    assert len(T.BLOCK_SHAPES) == 9
    assert all(w in (1, 2, 4) and r in (1, 2, 4) for w, r in T.BLOCK_SHAPES)


def test_pick_prefers_the_earlier_of_near_ties():
    assert T.pick({1: 1.0, 2: 0.9995}) == 1
    assert T.pick({1: 1.0, 2: 0.9}) == 2
    assert T.pick({(1, 1): 2.0, (2, 4): 1.5, (4, 4): 1.6}) == (2, 4)


def test_cache_key_fields():
    k = T.cache_key("gfx950:sramecc+:xnack-", 256, "k_ebe", 60, "mixed", 1, 1_000_000)
    assert k == "gfx950:sramecc+:xnack-|cu256|k_ebe|ntot60|mixed|k1|nels~2^20"


class _Ctx:
    ntot, nels, neq = 60, 1000, 10

    def __init__(self):
        self.shape = {"ebe_epl": 0, "blk_waves": 0, "blk_rows": 0}
        self.block_calls = 0

    def device_info(self):
        return "gfx950", 256

    def tune_get(self):
        return dict(self.shape)

    def tune_set(self, ebe_epl=0, blk_waves=0, blk_rows=0):
        self.shape = {"ebe_epl": ebe_epl, "blk_waves": blk_waves, "blk_rows": blk_rows}

    def block_apply(self, x, precision):
        pass

    def block_time_ebe(self, reps):
        self.block_calls += 1
        return 1.0 + abs(self.shape["blk_waves"] - 2) * 0.1     # fastest at 2


def test_block_tuning_and_cache(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("PARAFEM_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(T, "_memory", {})
    c = _Ctx()
    c.shape["blk_rows"] = 1                       # the FP64 passes are left as they are
    got = T.tune_block(c)
    assert got["blk_waves"] == 2 and not got["cached"]
    assert c.shape["blk_waves"] == 2 and c.shape["blk_rows"] == 1 and c.block_calls == 3
    data = json.loads(path.read_text())
    (key, entry), = data.items()
    assert "k_blk_ebe_sym" in key and entry["blk_waves"] == 2 and "blk_rows" not in entry
    monkeypatch.setattr(T, "_memory", {})
    c2 = _Ctx()
    again = T.tune_block(c2)
    assert again["cached"] and c2.block_calls == 0 and c2.shape["blk_waves"] == 2


def test_unwritable_cache_keeps_results_in_memory(tmp_path, monkeypatch):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    monkeypatch.setenv("PARAFEM_AMD_TUNE_CACHE", str(blocker / "tune.json"))
    monkeypatch.setattr(T, "_memory", {})
    c = _Ctx()
    T.tune_block(c)
    c2 = _Ctx()
    assert T.tune_block(c2)["cached"]
