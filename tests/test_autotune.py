"""Autotuner bookkeeping (no GPU): candidate sets and the tuning table."""
import json

from pelephysics_amd import autotune


def test_candidates():
    c = autotune.candidates("drm19", "ReactorBDFAdaptiveMixed", "1d")
    groups = [x for x in c if x["batched_cells_per_wave"]]
    c = [x for x in c if not x["batched_cells_per_wave"]]
    # group kernels: the table-driven drm19 only, every size, flat route
    assert {x["kernels"] for x in groups} == {"drm19_table"}
    assert {x["batched_cells_per_wave"] for x in groups} == set(autotune.GROUPS)
    assert len(groups) == len(autotune.GROUPS) * len(autotune.GROUP_WAVES) * len(autotune.GROUP_LU)
    assert not any(x["batched_box_direct"] for x in groups)
    got = {(x["kernels"], x["batched_waves_per_simd"], x["batched_waves_per_cu"], x["batched_lu_batch"],
            x["batched_lu_wait"]) for x in c}
    assert got == {(k, o, w, b, t) for k in ("drm19", "drm19_table") for o, w in autotune.OCCUPANCY
                   for b, t in autotune.LU_SERVICE}
    assert all(w <= 4 * o for _, o, w, _, _ in got)  # waves per CU within 4 SIMDs x waves per SIMD
    box = [x for x in autotune.candidates("drm19", "ReactorBDFAdaptiveMixed", "box")
           if not x["batched_cells_per_wave"]]
    one_wave = sum(1 for x in c if x["batched_waves_per_simd"] == 1)
    assert len(box) == 2 * len(c) - one_wave  # no box-kernel route for the 1-wave build
    assert not any(x["batched_waves_per_simd"] == 1 and x["batched_box_direct"] == 1 for x in box)
    c = autotune.candidates("dodecane_lu", "ReactorBDFAdaptive")
    assert c == [{"kernels": "dodecane_lu", "batched_waves_per_simd": 0, "batched_waves_per_cu": 0,
                  "batched_lu_batch": 0, "batched_lu_wait": 0, "batched_box_direct": 0}]
    # the plain adaptive BDF of a register-resident mechanism: three builds per implementation
    c = autotune.candidates("drm19", "ReactorBDFAdaptive", "box")
    assert {(x["kernels"], x["batched_waves_per_simd"]) for x in c} == {
        (k, o) for k in ("drm19", "drm19_table") for o in autotune.PLAIN_OCCUPANCY}


def test_lookup_reads_the_table(tmp_path, monkeypatch):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"drm19|ReactorBDFAdaptiveMixed|box": {"kernels": "drm19_table",
                                                                    "batched_waves_per_cu": 6}}))
    monkeypatch.setenv("PELE_CHEM_TUNING", str(p))
    e = autotune.lookup("drm19", "ReactorBDFAdaptiveMixed", "box")
    assert e["kernels"] == "drm19_table" and e["batched_waves_per_cu"] == 6
    assert autotune.lookup("drm19", "ReactorBDFAdaptiveMixed", "1d") is None
