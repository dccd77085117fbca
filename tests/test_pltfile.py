"""AMReX plotfile reading and the PeleLMeX reactor state
(Testing/Exec/ReactEval/main.cpp:310-387, PltFileManager.cpp:23-180).

The reference holds no plotfile, so the reader is checked against files of
the published AMReX format written here: multi-box round trips (bitwise),
both byte orders, uncovered cells, the MKS -> CGS conversion and component
ranges of the ReactEval driver, and its abort on a missing variable.  Against
real PeleLMeX output the reader is unpinned."""
import os

import numpy as np
import pytest

from pelephysics_amd import pltfile


def _field(nv, nz, ny, nx, seed=0):
    return np.random.default_rng(seed).standard_normal((nv, nz, ny, nx)) * 10.0 ** np.arange(nv)[:, None, None, None]


@pytest.mark.parametrize("big_endian", [False, True])
def test_round_trip_multi_box(tmp_path, big_endian):
    d = _field(3, 5, 6, 7)
    p = str(tmp_path / "plt00010")
    pltfile.write(p, ["density", "Temp", "rhoH"], d, prob_lo=(0.0, -1.0, 2.0), prob_hi=(0.7, 0.2, 2.5),
                  time=1.25e-4, max_grid_size=4, big_endian=big_endian)
    P = pltfile.PlotFile(p)
    assert P.title == "HyperCLaw-V1.1" and P.variables == ["density", "Temp", "rhoH"]
    assert P.time == 1.25e-4 and P.nlevels == 1 and P.domains == [([0, 0, 0], [6, 5, 4])]
    assert P.prob_lo == [0.0, -1.0, 2.0] and P.prob_hi == [0.7, 0.2, 2.5]
    assert np.allclose(P.dx[0], [0.1, 0.2, 0.1])
    with open(os.path.join(p, "Level_0", "Cell_H")) as f:
        assert "(8 0" in f.read()  # 2 x 2 x 2 boxes of at most 4 cells
    got = P.level(0)
    assert got.shape == d.shape and np.array_equal(got, d)


def test_uncovered_cells_are_nan(tmp_path):
    d = _field(1, 4, 4, 4)
    p = str(tmp_path / "plt")
    pltfile.write(p, ["x"], d, max_grid_size=2)
    # drop the last box from the FabArray header: its cells are then uncovered
    h = os.path.join(p, "Level_0", "Cell_H")
    lines = open(h).read().split("\n")
    lines[4] = "(7 0"
    del lines[5 + 7]  # the 8th box
    i = lines.index("8")
    lines[i] = "7"
    del lines[i + 8]
    open(h, "w").write("\n".join(lines))
    got = pltfile.PlotFile(p).level(0)
    assert np.isnan(got[0, 2:, 2:, 2:]).all()
    mask = np.ones(d.shape, bool)
    mask[0, 2:, 2:, 2:] = False
    assert np.array_equal(got[mask], d[mask])


def test_reactor_state_mks_to_cgs(tmp_path, models):
    m = models["drm19"]
    sp = [s.name for s in m.species]
    ns = len(sp)
    names = ["x_velocity", "density"] + [f"rhoY({s})" for s in sp] + ["rhoH", "Temp"] + \
            [f"F_rhoY({s})" for s in sp] + ["F_rhoH", "divu"]
    d = np.abs(_field(len(names), 3, 4, 5, seed=2)) + 1.0
    p = str(tmp_path / "plt_lmex")
    pltfile.write(p, names, d, max_grid_size=3)
    st = pltfile.reactor_state(pltfile.PlotFile(p), sp)
    assert st["rY"].shape == (ns, 3, 4, 5) and st["T"].shape == (1, 3, 4, 5)
    assert np.array_equal(st["rY"], d[2:2 + ns] * 1.0e-3)
    assert np.array_equal(st["rE"], d[2 + ns:3 + ns] * 10.0)
    assert np.array_equal(st["T"], d[3 + ns:4 + ns])
    assert np.array_equal(st["rYsrc"], d[4 + ns:4 + 2 * ns] * 1.0e-3)
    assert np.array_equal(st["rEsrc"], d[4 + 2 * ns:5 + 2 * ns] * 10.0)
    assert all(a.flags["C_CONTIGUOUS"] and a.dtype == np.float64 for a in st.values())


def test_reactor_state_missing_variable(tmp_path, models):
    sp = [s.name for s in models["drm19"].species]
    names = [f"rhoY({s})" for s in sp] + ["rhoH", "Temp"] + [f"F_rhoY({s})" for s in sp]  # no F_rhoH
    p = str(tmp_path / "plt_bad")
    pltfile.write(p, names, np.ones((len(names), 2, 2, 2)))
    with pytest.raises(pltfile.PltFileError, match="must be in the pltfile"):
        pltfile.reactor_state(pltfile.PlotFile(p), sp)
