"""The Manifold EOS's tabulated model on the host: the table file and the EOS
parameters that ``ReactorRK64.set_eos_parm`` hands to the library.

Reference: ``Source/Utility/BlackBoxFunction/Table.H`` (TabulatedFunctionData,
``read_table`` :63-139: the binary layout of a ``.ctb`` table),
``Source/Utility/BlackBoxFunction/BlackBoxFunction.H:11-53`` (64-character
names, ``get_var_index``) and ``Source/Eos/EosParams.H:39-160`` (EosParm<Manifold>
and how its indices are found from the ``manifold.*`` keys).

A ``.ctb`` file holds, in native byte order: Ndim (int32); Ndim dimension names
of 64 characters; Ndim grid lengths (int32); Nvar (int32); the Ndim grids
concatenated (float64); the model name (64 characters); Nvar variable names (64
characters each); the values, Nvar blocks of prod(lengths) float64 with the
first dimension fastest.
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass, field
from typing import List

import numpy as np

LEN_STR = 64  # BlackBoxFunctionData::len_str
MAXD = 5      # MAXD_TABLE
DENSITY_LOOKUP = {"linear": (0, "RHO"), "log": (1, "lnRHO"), "inverse": (2, "invRHO")}


NN_MAX_LAYERS = 16  # layers the device evaluates
NN_MAX_WIDTH = 100  # MAXD_NETWORK: every layer narrower than this
NN_KINDS = {"Linear": 0, "LeakyReLU": 1, "BatchNorm1d": 2}
# the layer order of a CMLM network (NeuralNetModelDef.H:170-205)
CMLM_ORDER = ["Linear", "LeakyReLU", "BatchNorm1d", "Linear", "LeakyReLU", "BatchNorm1d", "Linear"]


class PPManifold(C.Structure):
    """pp_chem_manifold (include/pele_chem_hip.h)."""
    _fields_ = [("ndim", C.c_int), ("dim_lengths", C.POINTER(C.c_int)), ("grids", C.POINTER(C.c_double)),
                ("nvar", C.c_int), ("values", C.POINTER(C.c_double)), ("idx_density", C.c_int),
                ("density_lookup", C.c_int), ("idx_T", C.c_int), ("compute_temperature", C.c_int),
                ("idx_src", C.c_int * MAXD), ("nominal_pressure", C.c_double),
                ("model", C.c_int), ("nn_layers", C.c_int), ("nn_kind", C.POINTER(C.c_int)),
                ("nn_nout", C.POINTER(C.c_int)), ("nn_params", C.POINTER(C.c_double))]


def _name(b: bytes) -> str:
    return b.decode("ascii", "replace").strip().strip("\x00").strip()


@dataclass
class TabulatedFunction:
    """TabulatedFunctionData: grids per dimension and the values of each
    variable, values[v] shaped (len[ndim-1], ..., len[0]) (first dimension
    fastest in memory)."""
    dimnames: List[str]
    grids: List[np.ndarray]
    varnames: List[str]
    values: np.ndarray  # (nvar, prod(lengths)) flat, first dimension fastest
    model_name: str = ""

    @property
    def ndim(self) -> int:
        return len(self.grids)

    @property
    def nvar(self) -> int:
        return len(self.varnames)

    @property
    def dim_lengths(self) -> List[int]:
        return [len(g) for g in self.grids]

    @classmethod
    def read(cls, path: str) -> "TabulatedFunction":
        """read_table (Table.H:63-139)."""
        with open(path, "rb") as f:
            b = f.read()
        o = 0

        def ints(n):
            nonlocal o
            v = struct.unpack_from(f"={n}i", b, o)
            o += 4 * n
            return list(v)

        def strs(n):
            nonlocal o
            v = [_name(b[o + i * LEN_STR:o + (i + 1) * LEN_STR]) for i in range(n)]
            o += n * LEN_STR
            return v

        def reals(n):
            nonlocal o
            v = np.frombuffer(b, dtype=np.float64, count=n, offset=o).copy()
            o += 8 * n
            return v

        (ndim,) = ints(1)
        if not 1 <= ndim <= MAXD:
            raise ValueError(f"{path}: {ndim} table dimensions (1..{MAXD} allowed)")
        dimnames = strs(ndim)
        lens = ints(ndim)
        (nvar,) = ints(1)
        flat = reals(sum(lens))
        grids, s = [], 0
        for n in lens:
            grids.append(flat[s:s + n])
            s += n
        (model,) = strs(1)
        varnames = strs(nvar)
        npts = int(np.prod(lens))
        values = reals(nvar * npts).reshape(nvar, npts)
        if o != len(b):
            raise ValueError(f"{path}: {len(b) - o} trailing bytes")
        return cls(dimnames, grids, varnames, values, model)

    def write(self, path: str) -> None:
        """The same layout (for tables built in tests and tools)."""
        def pad(s, fill=b" "):
            e = s.encode("ascii")
            if len(e) > LEN_STR:
                raise ValueError(f"name longer than {LEN_STR} characters: {s!r}")
            return e.ljust(LEN_STR, fill)
        out = [struct.pack("=i", self.ndim), b"".join(pad(n) for n in self.dimnames),
               struct.pack(f"={self.ndim}i", *self.dim_lengths), struct.pack("=i", self.nvar),
               np.concatenate(self.grids).astype(np.float64).tobytes(), pad(self.model_name, b"\0"),
               b"".join(pad(n) for n in self.varnames),
               np.ascontiguousarray(self.values, dtype=np.float64).tobytes()]
        with open(path, "wb") as f:
            f.write(b"".join(out))

    def var_index(self, name: str, require: bool = True) -> int:
        """get_var_index (BlackBoxFunction.H:32-53): -1 when absent and not required."""
        for i, v in enumerate(self.varnames):
            if v == name.strip():
                return i
        if require:
            raise KeyError(f"Requested variable not found in table: {name}")
        return -1


def _parse_fmt(fmt: str):
    """read_fmt (NeuralNetModelDef.H:17-81): "f(50,2)f(50)" -> [("f", [50, 2]), ("f", [50])]."""
    out, i = [], 0
    while i < len(fmt):
        t = fmt[i]
        if i + 1 >= len(fmt) or fmt[i + 1] != "(":
            raise ValueError(f"Error parsing neural net layer format. At character '{fmt[i]}' and index {i}.")
        j = fmt.index(")", i)
        out.append((t, [int(x) for x in fmt[i + 2:j].split(",")]))
        i = j + 1
    return out


@dataclass
class NeuralNetwork:
    """A home-rolled network of the Manifold EOS (``manifold.model =
    NeuralNet``): the ``.pnn`` weights file and its info file
    (NeuralNetHomerolled.H:13-200, NeuralNetModelDef.H:84-170,
    NeuralNetLayerDef.H).

    A ``.pnn`` file holds, in native byte order: the string length, the real
    size and the integer size in bytes and the input count (int32 each); then,
    to the end of the file, for each layer its name and format (one string
    each, NUL-padded) and its arrays: Linear a weight [nout][nin] and a bias
    [nout], LeakyReLU the negative slope, BatchNorm1d a scale [n] and a shift
    [n] (the normalisation folded into an affine map).  Weights and slopes
    have the real size, biases and shifts are read with the integer size, as
    the reference reads them.  The info file names the model, the input
    (``dimnames``) and output (``varnames``) variables."""
    model_name: str
    dimnames: List[str]
    varnames: List[str]
    nmanpar: int
    layers: list  # (kind name, nout, float64 parameters)

    @property
    def ndim(self) -> int:
        return len(self.dimnames)

    @property
    def nvar(self) -> int:
        return len(self.varnames)

    @staticmethod
    def read_info(path: str) -> dict:
        """read_metadata (NeuralNetHomerolled.H:95-185): ``key = values``,
        ``#`` comments, a trailing backslash continues the values."""
        umap, cur, cont = {}, None, False
        for line in open(path).read().splitlines():
            line = line.split("#", 1)[0]
            toks = line.split()
            if not toks:
                cont = False
                continue
            if not cont:
                if "=" not in toks:
                    continue
                k = toks.index("=")
                cur = toks[0].lower()
                toks = toks[k + 1:]
            cont = bool(toks) and toks[-1] == "\\"
            umap.setdefault(cur, []).extend(t for t in toks if t != "\\")
        return umap

    @classmethod
    def read(cls, pnn: str, info: str, cmlm_net: bool = True) -> "NeuralNetwork":
        if not pnn.strip().endswith(".pnn"):
            raise ValueError("Home-rolled neural net file should have .pnn extension.")
        if not cmlm_net:  # the device model exists only for CMLM nets (NeuralNetModelDef.H:207-209)
            raise ValueError("Only CMLM nets supported for now without owning data")
        raw = open(pnn, "rb").read()
        ssize, fsize, isize, inp = struct.unpack_from("4i", raw, 0)
        off = 16

        def reals(size, n):
            nonlocal off
            code = {4: "f", 8: "d"}.get(size)
            if code is None:
                raise ValueError(f"Cannot read data of size {size} bytes into array of amrex::Reals.")
            a = np.frombuffer(raw, dtype=np.dtype(code), count=n, offset=off).astype(np.float64)
            off += size * n
            return a

        layers, width = [], inp
        while off < len(raw):
            name = raw[off:off + ssize].split(b"\x00", 1)[0].decode().strip()
            fmt = raw[off + ssize:off + 2 * ssize].split(b"\x00", 1)[0].decode().strip()
            off += 2 * ssize
            f = _parse_fmt(fmt)
            if name == "Linear":
                (_, ws), (_, bs) = f
                if len(f) != 2 or len(ws) != 2 or ws[0] != bs[0]:
                    raise ValueError(f"bad Linear layer format {fmt}")
                w, b = reals(fsize, ws[0] * ws[1]), reals(isize, bs[0])
                if ws[1] != width:
                    raise ValueError(f"Linear layer takes {ws[1]} inputs, the previous layer gives {width}")
                width = ws[0]
                layers.append(("Linear", width, np.concatenate([w, b])))
            elif name == "LeakyReLU":
                layers.append(("LeakyReLU", width, reals(fsize, 1)))
            elif name == "BatchNorm1d":
                (_, ws), (_, bs) = f
                if ws[0] != bs[0] or ws[0] != width:
                    raise ValueError(f"bad BatchNorm1d layer format {fmt}")
                layers.append(("BatchNorm1d", width, np.concatenate([reals(fsize, ws[0]), reals(isize, bs[0])])))
            else:
                raise ValueError(f"unknown layer {name!r}")
        if [l[0] for l in layers] != CMLM_ORDER:
            raise ValueError("a CMLM net has the 7 layers Linear, LeakyReLU, BatchNorm1d, Linear, LeakyReLU, "
                             "BatchNorm1d, Linear")
        u = cls.read_info(info)
        for k in ("model_name", "ndim", "nvar", "nmanpar"):
            if len(u.get(k, [])) != 1:
                raise ValueError(f"info file: {k} needs one value")
        net = cls(u["model_name"][0], list(u.get("dimnames", [])), list(u.get("varnames", [])),
                  int(u["nmanpar"][0]), layers)
        if net.ndim != int(u["ndim"][0]) or net.nvar != int(u["nvar"][0]):
            raise ValueError("info file: dimnames / varnames do not match ndim / nvar")
        if net.ndim != inp or layers[-1][1] != net.nvar:
            raise ValueError("the network's input and output sizes differ from the info file's")
        if max(l[1] for l in layers) >= NN_MAX_WIDTH:
            raise ValueError("a layer is wider than MAXD_NETWORK")
        return net

    def var_index(self, name: str, require: bool = True) -> int:
        if name in self.varnames:
            return self.varnames.index(name)
        if require:
            raise ValueError(f"Variable not found: {name}")
        return -1


@dataclass
class ManifoldParm:
    """EosParm<Manifold> with a tabulated model or a network (``table`` is a
    TabulatedFunction or a NeuralNetwork; EosParams.H:39-49, initialised
    as InitParm<EosParm<Manifold>>::host_initialize, :84-160, does from the
    ``manifold.*`` keys)."""
    table: TabulatedFunction
    nominal_pressure_cgs: float
    compute_temperature: bool
    has_mani_src: bool = True
    density_lookup_type: str = "linear"
    idx_density: int = field(init=False)
    idx_T: int = field(init=False)
    idx_src: List[int] = field(init=False)
    identifier = "Manifold"

    def __post_init__(self):
        if self.density_lookup_type not in DENSITY_LOOKUP:
            raise ValueError("Invalid density lookup type supplied")
        self.idx_density = self.table.var_index(DENSITY_LOOKUP[self.density_lookup_type][1])
        self.idx_T = self.table.var_index("T")
        # the source of parameter "Y-<name>" is "SRC_<name>", of any other
        # parameter p "SRC_p"; a missing source is zero (EosParams.H:137-154)
        src = []
        for d in self.table.dimnames:
            key = "SRC_" + (d[2:] if d.startswith("Y-") else d)
            src.append(self.table.var_index(key, require=False) if self.has_mani_src else -1)
        self.idx_src = src

    @classmethod
    def from_file(cls, path: str, nominal_pressure_cgs: float, compute_temperature: bool,
                  has_mani_src: bool = True, density_lookup_type: str = "linear") -> "ManifoldParm":
        """manifold.model = Table with manifold.table.filename = path."""
        return cls(TabulatedFunction.read(path), nominal_pressure_cgs, compute_temperature, has_mani_src,
                   density_lookup_type)

    @classmethod
    def from_network(cls, pnn: str, info: str, nominal_pressure_cgs: float, compute_temperature: bool,
                     has_mani_src: bool = True, density_lookup_type: str = "linear",
                     cmlm_net: bool = True) -> "ManifoldParm":
        """manifold.model = NeuralNet with manifold.network.filename = pnn and
        manifold.network.info_filename = info."""
        return cls(NeuralNetwork.read(pnn, info, cmlm_net), nominal_pressure_cgs, compute_temperature,
                   has_mani_src, density_lookup_type)

    def c_struct(self):
        """(PPManifold, the arrays it points into: keep them alive for the call)."""
        t = self.table
        if isinstance(t, NeuralNetwork):
            kinds = (C.c_int * len(t.layers))(*[NN_KINDS[l[0]] for l in t.layers])
            nouts = (C.c_int * len(t.layers))(*[l[1] for l in t.layers])
            params = np.ascontiguousarray(np.concatenate([l[2] for l in t.layers]), dtype=np.float64)
            m = PPManifold()
            m.ndim, m.nvar = t.ndim, t.nvar
            m.model, m.nn_layers = 1, len(t.layers)
            m.nn_kind, m.nn_nout = kinds, nouts
            m.nn_params = params.ctypes.data_as(C.POINTER(C.c_double))
            self._fill_indices(m)
            return m, (kinds, nouts, params)
        lens = (C.c_int * t.ndim)(*t.dim_lengths)
        grids = np.ascontiguousarray(np.concatenate(t.grids), dtype=np.float64)
        values = np.ascontiguousarray(t.values, dtype=np.float64).reshape(-1)
        m = PPManifold()
        m.ndim = t.ndim
        m.dim_lengths = lens
        m.grids = grids.ctypes.data_as(C.POINTER(C.c_double))
        m.nvar = t.nvar
        m.values = values.ctypes.data_as(C.POINTER(C.c_double))
        self._fill_indices(m)
        return m, (lens, grids, values)

    def _fill_indices(self, m) -> None:
        m.idx_density = self.idx_density
        m.density_lookup = DENSITY_LOOKUP[self.density_lookup_type][0]
        m.idx_T = self.idx_T
        m.compute_temperature = 1 if self.compute_temperature else 0
        for d in range(MAXD):
            m.idx_src[d] = self.idx_src[d] if d < self.table.ndim else -1
        m.nominal_pressure = float(self.nominal_pressure_cgs)
