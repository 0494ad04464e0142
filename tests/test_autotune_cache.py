"""Host logic of the autotuner (no device): shape buckets and the decision
cache's JSON round trip."""
import json

from vectorta_amd import autotune as at


def test_shape_bucket_rounds_log2():
    assert at.shape_bucket(200_000, 1_250) == (18, 10)      # 2^17.6, 2^10.3
    assert at.shape_bucket(1, 0) == (0, 0)
    # shapes within a factor ~1.4 of one power of two share a decision
    assert at.shape_bucket(190_000) == at.shape_bucket(250_000)
    assert at.shape_bucket(100_000) != at.shape_bucket(300_000)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    t = at.Autotuner(path)
    assert t.lookup("gfx950", at.ALMA_MFMA_VARIANT, (200_000, 1_250)) is None
    t.record("gfx950", at.ALMA_MFMA_VARIANT, (200_000, 1_250), 3, {1: 0.30, 3: 0.27, 5: None})
    t2 = at.Autotuner(path)
    assert t2.lookup("gfx950", at.ALMA_MFMA_VARIANT, (210_000, 1_300)) == 3
    assert t2.lookup("gfx950", at.SCAN_WAVES, (210_000, 1_300)) is None
    assert t2.lookup("gfx942", at.ALMA_MFMA_VARIANT, (200_000, 1_250)) is None
    data = json.load(open(path))
    (entry,) = data.values()
    assert entry["ms"]["5"] is None and entry["value"] == 3


def test_candidates_are_in_range():
    # the ranges vta_ctx_set_tuning accepts (include/vta.h, vta_tune_knob)
    lo_hi = {at.ALMA_MFMA_VARIANT: (1, 11), at.SCAN_WAVES: (64, 65536), at.FIR_GROUP: (1, 256),
             at.MS_CHUNK: (64, 1 << 30), at.ROLL_FORM: (1, 2)}
    for knob, cands in at.CANDIDATES.items():
        lo, hi = lo_hi[knob]
        assert all(lo <= c <= hi for c in cands), knob
