"""AMReX plotfiles in and out, and the reactor state a PeleLMeX plotfile holds.

The reference's ReactEval driver can start from a PeleLMeX plotfile
(Testing/Exec/ReactEval/main.cpp:310-387): it reads the variables
"rhoY(<first species>)" .. (NUM_SPECIES consecutive components), "Temp",
"rhoH", "F_rhoY(<first species>)" .. and "F_rhoH" through
pele::physics::pltfilemanager::PltFileManager
(Source/Utility/PltFileManager/PltFileManager.cpp:23-180), aborts when one of
them is missing (main.cpp:336-340), and converts the MKS data to CGS: rhoY_k
and F_rhoY_k x 1e-3, rhoH and F_rhoH x 10 (main.cpp:364-383).

This module reads the plotfile format itself (AMReX "HyperCLaw-V1.1": a text
Header, per level a Cell_H FabArray header and Cell_D_* files of FABs, each a
text line naming its box, component count and real format followed by the
binary data, x fastest, component slowest) into numpy arrays in the layout of
an amrex::Array4 over the whole level domain: (ncomp, nz, ny, nx), C order.
PlotFile.level() reads a level at its own resolution (cells no box covers
are NaN); fill_patch() fills a level of another resolution from every level
of the file as PltFileManager::fillPatchFromPlt does (PltFileManager.cpp:
188-352).  write() produces plotfiles in the same format, with refined levels
if asked.
"""
from __future__ import annotations

import os
import re

import numpy as np

_BOX = re.compile(r"\(\(\s*([-\d,\s]+)\)\s*\(\s*([-\d,\s]+)\)\s*\(\s*([-\d,\s]+)\)\)")
_LE8 = "(8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1))"
_BE8 = "(8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8))"


class PltFileError(RuntimeError):
    pass


def _ints(s: str) -> list:
    return [int(x) for x in s.replace(",", " ").split()]


def _boxes(text: str) -> list:
    """Every ((lo) (hi) (type)) in text as (lo, hi) integer triples."""
    return [(_ints(m.group(1)), _ints(m.group(2))) for m in _BOX.finditer(text)]


def _dtype(fab_header: str) -> np.dtype:
    """The numpy type of a FAB's real descriptor: bytes per real and byte order."""
    m = re.search(r"\(\s*(\d+)\s*,\s*\(([\d\s]+)\)\s*\)\s*\)\s*\(\(", fab_header)
    if m is None:
        raise PltFileError(f"unreadable FAB header: {fab_header[:120]!r}")
    size, order = int(m.group(1)), _ints(m.group(2))
    if size not in (4, 8) or sorted(order) != list(range(1, size + 1)):
        raise PltFileError(f"unsupported real format in FAB header: {fab_header[:120]!r}")
    if order == list(range(size, 0, -1)):
        return np.dtype(f"<f{size}")
    if order == list(range(1, size + 1)):
        return np.dtype(f">f{size}")
    raise PltFileError(f"mixed byte order in FAB header: {fab_header[:120]!r}")


class PlotFile:
    """An AMReX plotfile directory: header metadata and level data."""

    def __init__(self, path: str):
        self.path = path
        with open(os.path.join(path, "Header")) as f:
            lines = f.read().split("\n")
        it = iter(lines)
        self.title = next(it).strip()
        self.nvars = int(next(it).split()[0])
        self.variables = [next(it).split()[0] for _ in range(self.nvars)]
        self.spacedim = int(next(it).split()[0])
        if self.spacedim != 3:
            raise PltFileError(f"{path}: {self.spacedim}-D plotfile, only 3-D is read")
        self.time = float(next(it).split()[0])
        self.nlevels = int(next(it).split()[0]) + 1
        self.prob_lo = [float(x) for x in next(it).split()]
        self.prob_hi = [float(x) for x in next(it).split()]
        self.ref_ratio = _ints(next(it))
        doms = _boxes(next(it))
        if len(doms) != self.nlevels:
            raise PltFileError(f"{path}: {len(doms)} level domains for {self.nlevels} levels")
        self.domains = doms
        self.level_steps = _ints(next(it))
        self.dx = [[float(x) for x in next(it).split()] for _ in range(self.nlevels)]
        self.coord_sys = int(next(it).split()[0])
        next(it)  # boundary width
        self.level_dirs = []
        for lev in range(self.nlevels):
            words = next(it).split()
            ngrids = int(words[1])
            next(it)  # level step
            for _ in range(ngrids * self.spacedim):
                next(it)  # physical extent of each grid
            self.level_dirs.append(next(it).strip())

    def index(self, name: str) -> int:
        try:
            return self.variables.index(name)
        except ValueError:
            return -1

    def domain_shape(self, lev: int = 0) -> tuple:
        """(nx, ny, nz) cells of level lev's domain."""
        lo, hi = self.domains[lev]
        return tuple(h - l + 1 for l, h in zip(lo, hi))

    def level(self, lev: int = 0) -> np.ndarray:
        """All variables of level lev over its domain, (nvars, nz, ny, nx)."""
        base = os.path.join(self.path, self.level_dirs[lev])
        with open(base + "_H") as f:
            text = f.read()
        head = text.split("\n")
        ncomp = int(head[2].split()[0])
        if ncomp != self.nvars:
            raise PltFileError(f"{base}_H: {ncomp} components, the Header lists {self.nvars}")
        ba_text = text[text.index("("):]
        nboxes = int(ba_text[1:].split()[0])
        boxes = _boxes(ba_text)[:nboxes]
        fods = re.findall(r"FabOnDisk:\s+(\S+)\s+(\d+)", text)
        if len(boxes) != nboxes or len(fods) != nboxes:
            raise PltFileError(f"{base}_H: {nboxes} boxes, {len(boxes)} read, {len(fods)} FabOnDisk entries")
        (dlo, dhi) = self.domains[lev]
        shape = [h - l + 1 for l, h in zip(dlo, dhi)]
        out = np.full((ncomp, shape[2], shape[1], shape[0]), np.nan)
        lvdir = os.path.dirname(base)
        for (lo, hi), (fname, off) in zip(boxes, fods):
            with open(os.path.join(lvdir, fname), "rb") as f:
                f.seek(int(off))
                fab_header = f.readline().decode("ascii")
                fb = _boxes(fab_header)
                if not fb or fb[0] != (lo, hi):
                    raise PltFileError(f"{fname}@{off}: FAB box {fb[:1]} is not the Cell_H box {(lo, hi)}")
                nc = int(fab_header.rsplit(")", 1)[1].split()[0])
                dt = _dtype(fab_header)
                n = [h - l + 1 for l, h in zip(lo, hi)]
                count = nc * n[0] * n[1] * n[2]
                raw = f.read(count * dt.itemsize)
            if len(raw) != count * dt.itemsize:
                raise PltFileError(f"{fname}@{off}: FAB data truncated")
            data = np.frombuffer(raw, dtype=dt)
            sl = tuple(slice(l - d, h - d + 1) for l, h, d in zip(lo, hi, dlo))
            out[:nc, sl[2], sl[1], sl[0]] = data.astype(np.float64).reshape(nc, n[2], n[1], n[0])
        return out


def write(path: str, names, data: np.ndarray, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0),
          time: float = 0.0, max_grid_size: int = 32, big_endian: bool = False, fine=()) -> None:
    """A plotfile of data (nvars, nz, ny, nx) on the level-0 domain
    (0,0,0)-(nx-1,ny-1,nz-1), split into boxes of at most max_grid_size cells
    per direction, with refined levels from `fine`: a sequence of (ratio to
    the level below, boxes [((lo), (hi)), ...] in that level's index space,
    data (nvars, nz_l, ny_l, nx_l) over the whole level domain, read only
    inside the boxes).  Each level's FABs go to Level_<l>/Cell_D_00000."""
    data = np.asarray(data, dtype=np.float64)
    nv, nz, ny, nx = data.shape
    if len(names) != nv:
        raise ValueError(f"{len(names)} names for {nv} variables")
    levels = [((nx, ny, nz), [((i, j, k), (min(i + max_grid_size, nx) - 1, min(j + max_grid_size, ny) - 1,
                                           min(k + max_grid_size, nz) - 1))
                              for k in range(0, nz, max_grid_size) for j in range(0, ny, max_grid_size)
                              for i in range(0, nx, max_grid_size)], data)]
    ratios = []
    for r, boxes, d in fine:
        n0 = levels[-1][0]
        shape = tuple(s * r for s in n0)
        d = np.asarray(d, dtype=np.float64)
        if d.shape != (nv, shape[2], shape[1], shape[0]):
            raise ValueError(f"level {len(levels)}: data shape {d.shape}, expected {(nv,) + shape[::-1]}")
        levels.append((shape, [(tuple(lo), tuple(hi)) for lo, hi in boxes], d))
        ratios.append(r)

    def bstr(lo, hi):
        return f"(({lo[0]},{lo[1]},{lo[2]}) ({hi[0]},{hi[1]},{hi[2]}) (0,0,0))"

    os.makedirs(path, exist_ok=True)
    nlev = len(levels)
    dxs = [[(prob_hi[d] - prob_lo[d]) / s for d, s in enumerate(shape)] for shape, _, _ in levels]
    hdr = ["HyperCLaw-V1.1", str(nv)] + list(names) + [
        "3", repr(float(time)), str(nlev - 1), " ".join(repr(float(x)) for x in prob_lo),
        " ".join(repr(float(x)) for x in prob_hi), " ".join(str(r) for r in ratios),
        " ".join(bstr((0, 0, 0), tuple(s - 1 for s in shape)) for shape, _, _ in levels),
        " ".join("0" for _ in levels)]
    hdr += [" ".join(repr(x) for x in dx) for dx in dxs]
    hdr += ["0", "0"]
    for lev, (shape, boxes, _) in enumerate(levels):
        hdr.append(f"{lev} {len(boxes)} {float(time)!r}")
        hdr.append("0")
        for lo, hi in boxes:
            for d in range(3):
                hdr.append(f"{prob_lo[d] + lo[d] * dxs[lev][d]!r} {prob_lo[d] + (hi[d] + 1) * dxs[lev][d]!r}")
        hdr.append(f"Level_{lev}/Cell")
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(hdr) + "\n")
    fmt = _BE8 if big_endian else _LE8
    dt = np.dtype(">f8" if big_endian else "<f8")
    for lev, (shape, boxes, d) in enumerate(levels):
        os.makedirs(os.path.join(path, f"Level_{lev}"), exist_ok=True)
        offsets, mins, maxs = [], [], []
        with open(os.path.join(path, f"Level_{lev}", "Cell_D_00000"), "wb") as f:
            for lo, hi in boxes:
                offsets.append(f.tell())
                f.write(f"FAB ({fmt}){bstr(lo, hi)} {nv}\n".encode("ascii"))
                blk = d[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
                f.write(np.ascontiguousarray(blk).astype(dt).tobytes())
                mins.append(blk.reshape(nv, -1).min(1))
                maxs.append(blk.reshape(nv, -1).max(1))
        cell_h = ["1", "0", str(nv), "0", f"({len(boxes)} 0"] + [bstr(lo, hi) for lo, hi in boxes] + [
            ")", str(len(boxes))] + [f"FabOnDisk: Cell_D_00000 {o}" for o in offsets] + [
            "", f"{len(boxes)},{nv}"] + [",".join(repr(float(v)) for v in m) + "," for m in mins] + [
            "", f"{len(boxes)},{nv}"] + [",".join(repr(float(v)) for v in m) + "," for m in maxs]
        with open(os.path.join(path, f"Level_{lev}", "Cell_H"), "w") as f:
            f.write("\n".join(cell_h) + "\n")


def _interp(c: np.ndarray, rr, kind: str, periodic) -> np.ndarray:
    """Fine data from coarse data c (ncomp, nz, ny, nx) refined by rr = (rx,
    ry, rz): piecewise constant ("pc", PltFileManager's interp_type != 1) or
    cell-conservative linear ("cons", interp_type 1: per direction the
    monotonized-central slope min(|u+ - u-| / 2, 2 |u+ - u|, 2 |u - u-|) with
    the central difference's sign, 0 where the one-sided differences differ
    in sign; fine cell m of r at u + slope ((m + 1/2) / r - 1/2), so the fine
    cells of a coarse cell average to it).  Neighbours beyond a non-periodic
    domain boundary and cells without data (NaN) count as copies of the cell
    itself (first-order extrapolation): zero slope there."""
    out = c
    slopes = []
    for ax, r in ((3, rr[0]), (2, rr[1]), (1, rr[2])):
        if kind != "cons" or r == 1:
            slopes.append(None)
            continue
        per = periodic[3 - ax]
        up = np.roll(c, -1, axis=ax)
        dn = np.roll(c, 1, axis=ax)
        if not per:
            idx = [slice(None)] * 4
            idx[ax] = -1
            up[tuple(idx)] = c[tuple(idx)]
            idx[ax] = 0
            dn[tuple(idx)] = c[tuple(idx)]
        up = np.where(np.isnan(up), c, up)
        dn = np.where(np.isnan(dn), c, dn)
        dc = 0.5 * (up - dn)
        df, db = 2.0 * (up - c), 2.0 * (c - dn)
        s = np.where(df * db > 0.0, np.sign(dc) * np.minimum(np.abs(dc), np.minimum(np.abs(df), np.abs(db))), 0.0)
        slopes.append(s)
    out = c
    for (ax, r), s in zip(((3, rr[0]), (2, rr[1]), (1, rr[2])), slopes):
        out = np.repeat(out, r, axis=ax)
    for (ax, r), s in zip(((3, rr[0]), (2, rr[1]), (1, rr[2])), slopes):
        if s is None:
            continue
        off = (np.arange(r) + 0.5) / r - 0.5
        reps = [1, 1, 1, 1]
        reps[ax] = out.shape[ax] // r
        shape = [1, 1, 1, 1]
        shape[ax] = out.shape[ax]
        o = np.tile(off, reps[ax]).reshape(shape)
        sf = s
        for (ax2, r2) in ((3, rr[0]), (2, rr[1]), (1, rr[2])):
            sf = np.repeat(sf, r2, axis=ax2)
        out = out + sf * o
    return out


def fill_patch(plt: PlotFile, shape, comps=None, interp_type: int = 1, periodic=(False, False, False)):
    """PltFileManager::fillPatchFromPlt (PltFileManager.cpp:188-352) for a
    level of `shape` = (nx, ny, nz) cells over the file's problem domain:
    every cell from level 0 of the file (copied, or interpolated when the
    level is finer: cell-conservative linear for interp_type 1, piecewise
    constant otherwise), then, level by level, the cells each finer level of
    the file covers (interpolated the same way, or copied from the level of
    the same resolution, after which the finer levels are not read).  A level
    coarser than the file's level 0 raises, as the reference aborts; file
    levels finer than `shape` are not used.  Returns (ncomp, nz, ny, nx)."""
    comps = list(range(plt.nvars)) if comps is None else list(comps)
    kind = "cons" if interp_type == 1 else "pc"

    def ratio(lev):
        s = plt.domain_shape(lev)
        if any(a % b or a < b for a, b in zip(shape, s)):
            return None
        return tuple(a // b for a, b in zip(shape, s))

    r0 = ratio(0)
    if r0 is None:
        raise PltFileError("When initializing dataFromPlt, our level 0 can't be coarser than the PltFile one")
    d0 = plt.level(0)[comps]
    out = d0 if r0 == (1, 1, 1) else _interp(d0, r0, kind, periodic)
    for lev in range(1, plt.nlevels):
        rr = ratio(lev)
        if rr is None:
            break
        dl = plt.level(lev)[comps]
        fine = dl if rr == (1, 1, 1) else _interp(dl, rr, kind, periodic)
        out = np.where(np.isnan(fine), out, fine)
        if rr == (1, 1, 1):
            break
    return np.ascontiguousarray(out)


def reactor_state(plt: PlotFile, species, lev: int = 0) -> dict:
    """The react() inputs of a PeleLMeX plotfile level, converted to CGS as
    ReactEval does (main.cpp:314-383): rY and rYsrc (N, nz, ny, nx), T, rE
    (rhoH) and rEsrc (1, nz, ny, nx), contiguous float64 in Array4 layout.
    The species are the N consecutive components from "rhoY(<species[0]>)"
    and from "F_rhoY(<species[0]>)"."""
    ns = len(species)
    idx = {k: plt.index(v) for k, v in (("rY", f"rhoY({species[0]})"), ("rYsrc", f"F_rhoY({species[0]})"),
                                        ("rE", "rhoH"), ("T", "Temp"), ("rEsrc", "F_rhoH"))}
    if min(idx.values()) < 0:
        raise PltFileError("All of 'rhoY(<spec>)', 'F_rhoY(<spec>)', 'rhoH', 'Temp' and 'F_rhoH' "
                           "must be in the pltfile !")
    for k in ("rY", "rYsrc"):
        if idx[k] + ns > plt.nvars:
            raise PltFileError(f"{plt.path}: {ns} species from component {idx[k]} exceed {plt.nvars} variables")
    d = plt.level(lev)
    return {
        "rY": np.ascontiguousarray(d[idx["rY"]:idx["rY"] + ns] * 1.0e-3),
        "rYsrc": np.ascontiguousarray(d[idx["rYsrc"]:idx["rYsrc"] + ns] * 1.0e-3),
        "T": np.ascontiguousarray(d[idx["T"]:idx["T"] + 1]),
        "rE": np.ascontiguousarray(d[idx["rE"]:idx["rE"] + 1] * 10.0),
        "rEsrc": np.ascontiguousarray(d[idx["rEsrc"]:idx["rEsrc"] + 1] * 10.0),
    }
