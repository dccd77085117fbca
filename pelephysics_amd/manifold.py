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


class PPManifold(C.Structure):
    """pp_chem_manifold (include/pele_chem_hip.h)."""
    _fields_ = [("ndim", C.c_int), ("dim_lengths", C.POINTER(C.c_int)), ("grids", C.POINTER(C.c_double)),
                ("nvar", C.c_int), ("values", C.POINTER(C.c_double)), ("idx_density", C.c_int),
                ("density_lookup", C.c_int), ("idx_T", C.c_int), ("compute_temperature", C.c_int),
                ("idx_src", C.c_int * MAXD), ("nominal_pressure", C.c_double)]


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


@dataclass
class ManifoldParm:
    """EosParm<Manifold> with a tabulated model (EosParams.H:39-49, initialised
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

    def c_struct(self):
        """(PPManifold, the arrays it points into: keep them alive for the call)."""
        t = self.table
        lens = (C.c_int * t.ndim)(*t.dim_lengths)
        grids = np.ascontiguousarray(np.concatenate(t.grids), dtype=np.float64)
        values = np.ascontiguousarray(t.values, dtype=np.float64).reshape(-1)
        m = PPManifold()
        m.ndim = t.ndim
        m.dim_lengths = lens
        m.grids = grids.ctypes.data_as(C.POINTER(C.c_double))
        m.nvar = t.nvar
        m.values = values.ctypes.data_as(C.POINTER(C.c_double))
        m.idx_density = self.idx_density
        m.density_lookup = DENSITY_LOOKUP[self.density_lookup_type][0]
        m.idx_T = self.idx_T
        m.compute_temperature = 1 if self.compute_temperature else 0
        for d in range(MAXD):
            m.idx_src[d] = self.idx_src[d] if d < t.ndim else -1
        m.nominal_pressure = float(self.nominal_pressure_cgs)
        return m, (lens, grids, values)
