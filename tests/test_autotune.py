"""Autotuner bookkeeping (no GPU): candidate sets and the tuning table."""
import json

from pelephysics_amd import autotune


def test_candidates():
    c = autotune.candidates("drm19", "ReactorBDFAdaptiveMixed", "1d")
    assert {(x["kernels"], x["batched_waves_per_cu"], x["batched_lu_batch"], x["batched_lu_wait"]) for x in c} == {
        (k, w, b, t) for k in ("drm19", "drm19_table") for w in autotune.WAVES for b, t in autotune.LU_SERVICE}
    assert len(autotune.candidates("drm19", "ReactorBDFAdaptiveMixed", "box")) == 2 * len(c)
    c = autotune.candidates("dodecane_lu", "ReactorBDFAdaptive")
    assert c == [{"kernels": "dodecane_lu", "batched_waves_per_cu": 0, "batched_lu_batch": 0,
                  "batched_lu_wait": 0, "batched_box_direct": 0}]


def test_lookup_reads_the_table(tmp_path, monkeypatch):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"drm19|ReactorBDFAdaptiveMixed|box": {"kernels": "drm19_table",
                                                                    "batched_waves_per_cu": 6}}))
    monkeypatch.setenv("PELE_CHEM_TUNING", str(p))
    e = autotune.lookup("drm19", "ReactorBDFAdaptiveMixed", "box")
    assert e["kernels"] == "drm19_table" and e["batched_waves_per_cu"] == 6
    assert autotune.lookup("drm19", "ReactorBDFAdaptiveMixed", "1d") is None
