"""The Python mirror checks every array before it hands a raw address to the
library: a tensor of another dtype, on the host, or of the wrong size would
otherwise be read and written out of bounds by the kernels.  These checks run
before any library call, so they are tested without a GPU."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pelephysics_amd.reactor import ReactorBase  # noqa: E402


def _reactor(ns=9):
    R = ReactorBase.create("ReactorBDF")
    R.nspec, R.mechanism = ns, "LiDryer"  # as init() sets them; no handle is created
    return R


def test_host_arrays_sizes_and_dtype():
    R = _reactor()
    nc = 3
    good = [np.zeros(nc * 10), np.zeros(nc * 9), np.zeros(nc), np.zeros(nc)]
    bad_size = [np.zeros(nc * 10 - 1)] + good[1:]
    with pytest.raises(ValueError, match="y: expected 30 elements"):
        R.react_1d(*bad_size, 1e-5)
    bad_src = good[:1] + [np.zeros(nc * 10)] + good[2:]
    with pytest.raises(ValueError, match="ysrc: expected 27 elements"):
        R.react_1d(*bad_src, 1e-5)
    f32 = [good[0].astype(np.float32)] + good[1:]
    with pytest.raises(ValueError, match="contiguous float64"):
        R.react_1d(*f32, 1e-5)
    mixed = good[:3] + [torch.zeros(nc, dtype=torch.float64)]
    with pytest.raises(ValueError):
        R.react_1d(*mixed, 1e-5)


def test_device_tensors_checked():
    R = _reactor()
    nc = 2
    cpu = [torch.zeros(nc * 10, dtype=torch.float64), torch.zeros(nc * 9, dtype=torch.float64),
           torch.zeros(nc, dtype=torch.float64), torch.zeros(nc, dtype=torch.float64)]
    with pytest.raises(ValueError, match="expected a GPU tensor"):
        R.react_1d(*cpu, 1e-5)
    with pytest.raises(ValueError, match="expected float64"):
        R.react_1d(cpu[0].float(), *cpu[1:], 1e-5)
    with pytest.raises(ValueError):
        R.react_1d_chained(*cpu, 1e-5, 3)
    with pytest.raises(ValueError, match="expected a GPU tensor"):
        R.adjoint_1d(cpu[0], cpu[2], 2, 1e-5)


def test_box_views_checked():
    R = _reactor()
    box = (2, 3, 4)
    rY = torch.zeros((9,) + box, dtype=torch.float64)
    T = torch.zeros(box, dtype=torch.float64)
    with pytest.raises(ValueError, match="expected a GPU tensor"):
        R.react(rY, rY, T, T, T, T, 1e-5)
    with pytest.raises(ValueError, match="expected float64"):
        R.react(rY, rY, T.float(), T, T, T, 1e-5)
    R._box_tensors  # the checker exists on every reactor
    with pytest.raises(ValueError):
        R.production_rate(torch.zeros((4, 9), dtype=torch.float64), torch.zeros(4, dtype=torch.float64))
