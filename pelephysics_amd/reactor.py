"""Python mirror of ``pele::physics::reactions::ReactorBase`` over the C-ABI.

Reference interface (Source/Reactions/ReactorBase.H:14-106, Source/Factory.H):

* ``ReactorBase.create(name)`` selects an integrator by its identifier
  ("ReactorBDF", "ReactorRK64", "ReactorCvode": ReactorCvode's options on
  this project's FP64 adaptive BDF, "ReactorArkode": its explicit RK tables
  and step controllers) plus this project's adaptive forms
  ("ReactorBDFAdaptive", "ReactorBDFAdaptiveMixed", "ReactorBDFBatchedFP64");
  unknown names raise with
  the list of valid names (``Factory.H:135-147`` aborts with the same list);
* ``init(reactor_type, ncells)`` takes the ``ode.*`` keys as keyword
  arguments (the ParmParse queries of ``ReactorBDF.cpp:118-128``,
  ``ReactorRK64.cpp:12-18``) and the mechanism name;
* ``react(box...)`` / ``react_1d(...)`` have the argument meaning of the two
  ``react`` overloads; arrays are torch tensors on the GPU (updated in place)
  or, for the 1-D form, host numpy arrays (copied in and out as the
  reference's GPU build does);
* the return value is ``int(average steps)`` (ReactorCvode: average RHS
  evaluations); errors raise ``ReactorError`` (the reference aborts).

The computation is done by ``libpele_chem_hip.so`` (HIP kernels for gfx950);
there is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; PELE_CHEM_HIP_LIB selects another build of the same library
LIB_PATH = os.environ.get("PELE_CHEM_HIP_LIB") or os.path.join(HERE, "libpele_chem_hip.so")

BDF, RK64, BDF_ADAPTIVE, BDF_ADAPTIVE_MIXED, BDF_BATCHED_FP64, CVODE, ARKODE = 0, 1, 2, 3, 4, 5, 6
E_REACTOR, H_REACTOR = 1, 2  # Source/Reactions/ReactorTypes.H:10-11

IDENTIFIERS = {
    "ReactorBDF": BDF,
    "ReactorRK64": RK64,
    "ReactorBDFAdaptive": BDF_ADAPTIVE,
    "ReactorBDFAdaptiveMixed": BDF_ADAPTIVE_MIXED,
    "ReactorBDFBatchedFP64": BDF_BATCHED_FP64,
    "ReactorCvode": CVODE,
    "ReactorArkode": ARKODE,
}

ERRORS = {-1: "EINVAL", -2: "EMECH", -3: "EHIP", -4: "EINTEGRATE"}


class ReactorError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("mechanism", C.c_char_p), ("reactor_type", C.c_int), ("mode", C.c_int),
        ("mod_reactor", C.c_int),
        ("bdf_nsubsteps", C.c_int), ("bdf_nonlinear_iters", C.c_int),
        ("bdf_gmres_restarts", C.c_int), ("bdf_gmres_kspiters", C.c_int),
        ("bdf_gmres_precond", C.c_int), ("bdf_scheme", C.c_int),
        ("bdf_nonlinear_tol", C.c_double), ("bdf_gmres_tol", C.c_double),
        ("rk64_atol", C.c_double), ("rk64_nsubsteps_guess", C.c_int),
        ("rk64_nsubsteps_min", C.c_int), ("rk64_nsubsteps_max", C.c_int),
        ("rtol", C.c_double), ("atol", C.c_double), ("first_step", C.c_double),
        ("max_steps", C.c_int), ("block_size", C.c_int), ("workspace_bytes", C.c_longlong),
        ("clean_init_massfrac", C.c_int), ("batched_waves_per_cu", C.c_int),
        ("batched_lu_batch", C.c_int), ("batched_lu_wait", C.c_int), ("batched_box_direct", C.c_int),
        ("batched_waves_per_simd", C.c_int), ("batched_cells_per_wave", C.c_int),
        ("cvode_max_order", C.c_int), ("max_nls_iters", C.c_int), ("jac_max_age", C.c_int),
        ("manifold_dim", C.c_int), ("erk_method", C.c_int), ("erk_controller", C.c_int),
    ]


class Array4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("jstride", C.c_longlong), ("kstride", C.c_longlong),
                ("nstride", C.c_longlong), ("begin", C.c_int * 3), ("end", C.c_int * 3),
                ("ncomp", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("ncells", C.c_longlong), ("nsteps", C.c_longlong), ("nrhs", C.c_longlong),
                ("njac", C.c_longlong), ("nlu", C.c_longlong), ("nnewton", C.c_longlong),
                ("nreject", C.c_longlong), ("max_steps_cell", C.c_int),
                ("min_steps_cell", C.c_int), ("nfailed", C.c_int), ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# (ode.* key, config field) pairs accepted by init()
_KEYS = {
    "bdf_nsubsteps": "bdf_nsubsteps", "bdf_nonlinear_iters": "bdf_nonlinear_iters",
    "bdf_gmres_restarts": "bdf_gmres_restarts", "bdf_gmres_kspiters": "bdf_gmres_kspiters",
    "bdf_gmres_precond": "bdf_gmres_precond", "bdf_scheme": "bdf_scheme",
    "bdf_nonlinear_tol": "bdf_nonlinear_tol", "bdf_gmres_tol": "bdf_gmres_tol",
    "atol": None,  # RK64 absolute tolerance and adaptive atol
    "rtol": "rtol", "first_step": "first_step", "max_steps": "max_steps",
    "rk64_nsubsteps_guess": "rk64_nsubsteps_guess", "rk64_nsubsteps_min": "rk64_nsubsteps_min",
    "rk64_nsubsteps_max": "rk64_nsubsteps_max", "mod_reactor": "mod_reactor",
    "block_size": "block_size", "workspace_bytes": "workspace_bytes",
    "clean_init_massfrac": "clean_init_massfrac", "batched_waves_per_cu": "batched_waves_per_cu",
    "batched_lu_batch": "batched_lu_batch", "batched_lu_wait": "batched_lu_wait",
    "batched_box_direct": "batched_box_direct", "batched_waves_per_simd": "batched_waves_per_simd",
    "batched_cells_per_wave": "batched_cells_per_wave",
    # ReactorCvode (and the plain adaptive BDF): cvode.max_order,
    # cvode.max_substeps, ode.max_nls_iters, the Jacobian's largest age
    "max_order": "cvode_max_order", "max_substeps": "max_steps", "max_nls_iters": "max_nls_iters",
    "jac_max_age": "jac_max_age",
    # mechanism "Manifold": MANIFOLD_DIM, the number of manifold parameters
    "manifold_dim": "manifold_dim",
    # ReactorArkode (ReactorArkode.cpp:12-18): ode.rk_method, ode.rk_controller;
    # ode.use_erkstep selects ERKStep or ARKStep there, both explicit here
    "rk_method": "erk_method", "rk_controller": "erk_controller", "use_erkstep": None,
}

_lib = None


def library() -> C.CDLL:
    """Load libpele_chem_hip.so (raises if it has not been built).

    torch is imported first: its HIP runtime is then the one the library
    binds to, so the tensors handed to the library and the library's own
    allocations live in one runtime (loaded the other way round, torch finds
    no GPU)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run python -m pelephysics_amd.build")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pp_chem_config_init.argtypes = [C.POINTER(Config), C.c_char_p, C.c_int, C.c_int]
        L.pp_chem_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.pp_chem_destroy.argtypes = [vp]
        L.pp_chem_set_typ_vals.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
        L.pp_chem_num_species.argtypes = [vp]
        L.pp_chem_num_reactions.argtypes = [vp]
        L.pp_chem_workspace_rows.argtypes = [vp]
        L.pp_chem_workspace_rows.restype = C.c_longlong
        L.pp_chem_adjoint_1d.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, vp, vp,
                                         vp]
        L.pp_chem_list_mechanisms.argtypes = [C.c_char_p, C.c_int]
        a4 = C.POINTER(Array4)
        L.pp_chem_react_box.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4,
                                         a4, a4, a4, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), vp]
        L.pp_chem_react_1d.argtypes = [vp, vp, vp, vp, vp, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.c_int, C.c_int, vp]
        L.pp_chem_flatten.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4, a4,
                                      vp, vp, vp, vp, vp]
        L.pp_chem_unflatten.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4, a4,
                                        vp, vp, vp, C.c_double, vp]
        L.pp_chem_react_1d_chained.argtypes = [vp, vp, vp, vp, vp, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.c_int, C.c_int, vp]
        L.pp_chem_react_1d_noalloc.argtypes = [vp, vp, vp, vp, vp, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.c_int, vp]
        L.pp_chem_production_rate.argtypes = [vp, vp, vp, C.c_int, vp, vp]
        L.pp_chem_ajacobian.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp]
        L.pp_chem_progress_rate.argtypes = [vp, vp, vp, C.c_int, vp, vp]
        L.pp_chem_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.pp_chem_copy_cell_steps.argtypes = [vp, vp, C.c_int]
        from .manifold import PPManifold
        L.pp_chem_set_manifold.argtypes = [vp, C.POINTER(PPManifold)]
        L.pp_chem_last_error.argtypes = [vp]
        L.pp_chem_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def mechanisms() -> Sequence[str]:
    L = library()
    n = L.pp_chem_list_mechanisms(None, 0)
    buf = C.create_string_buffer(n)
    L.pp_chem_list_mechanisms(buf, n)
    return buf.value.decode().split(",")


def _device_f64(t, name: str, numel: Optional[int] = None):
    """Raise ValueError unless t is a float64 tensor on a GPU (and, when
    numel is given, holds exactly that many elements).  The library takes raw
    device addresses, so a tensor of another dtype, a host tensor or a short
    buffer would be read and written out of bounds by the kernels."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != torch.float64:
        raise ValueError(f"{name}: expected float64, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: expected a GPU tensor, got device {t.device}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")
    return t


def _array4_from_tensor(t, ncomp_axis_first=True) -> Array4:
    """Array4 view of a contiguous torch tensor shaped (ncomp, nz, ny, nx) or (nz, ny, nx)."""
    if t.dim() == 3:
        t = t.unsqueeze(0)
    if t.dim() != 4:
        raise ValueError("expected (ncomp, nz, ny, nx)")
    nc, nz, ny, nx = t.shape
    st = t.stride()
    a = Array4()
    a.p = t.data_ptr()
    a.nstride, a.kstride, a.jstride = st[0], st[1], st[2]
    if st[3] != 1:
        raise ValueError("the i index must be unit-stride")
    a.begin[:] = [0, 0, 0]
    a.end[:] = [nx, ny, nz]
    a.ncomp = nc
    return a


class ReactorBase:
    """Factory + common driver; subclasses fix the integrator."""

    identifier = "ReactorBase"
    _mode: int = -1

    @staticmethod
    def create(name: str) -> "ReactorBase":
        cls = _REGISTRY.get(name)
        if cls is None:
            raise ReactorError(-1, f"unknown reactor {name!r}; valid: {sorted(_REGISTRY)}")
        return cls()

    def __init__(self):
        self._h = C.c_void_p()
        self.reactor_type = 0
        self.mechanism = None

    def init(self, reactor_type: int, ncells: int = 0, mechanism: str = "drm19", tuned: Optional[str] = None,
             **ode) -> int:
        """ReactorX::init(reactor_type, ncells); ode.* keys as keyword arguments.
        tuned = "box" or "1d": take the kernel implementation and launch shape
        the autotuner chose for this mechanism, integrator and form
        (pelephysics_amd.autotune), when it has an entry."""
        L = library()
        if tuned:
            from .autotune import lookup
            entry = lookup(mechanism, self.identifier, tuned)
            if entry:
                mechanism = entry["kernels"]
                for k in ("batched_waves_per_cu", "batched_lu_batch", "batched_lu_wait", "batched_box_direct",
                          "batched_waves_per_simd", "batched_cells_per_wave"):
                    if k in entry:
                        ode.setdefault(k, entry[k])
        cfg = Config()
        rc = L.pp_chem_config_init(C.byref(cfg), mechanism.encode(), self._mode, reactor_type)
        if rc != 0:
            raise ReactorError(rc, f"mechanism {mechanism!r} not compiled; have {mechanisms()}")
        for k, v in ode.items():
            if k not in _KEYS:
                raise ReactorError(-1, f"unknown ode key {k!r}")
            if k == "atol":
                cfg.rk64_atol = float(v)
                cfg.atol = float(v)
            elif k == "use_erkstep":
                continue
            else:
                setattr(cfg, _KEYS[k], v)
        self._cfg = cfg
        self.close()
        rc = L.pp_chem_create(C.byref(cfg), C.byref(self._h))
        if rc != 0:
            raise ReactorError(rc, "Wrong reactor type. Choose between 1 (e) or 2 (h)."
                               if reactor_type not in (1, 2) else "invalid configuration")
        self.reactor_type = reactor_type
        self.mechanism = mechanism
        self.nspec = L.pp_chem_num_species(self._h)
        return 0

    def set_typ_vals_ode(self, typ_vals):
        """ReactorBase::set_typ_vals_ode (ReactorBase.cpp:6-31): N+1 typical
        values (species, then T).  ReactorCvode then takes atol x typ_vals[k]
        as the absolute tolerance of component k (ReactorUtils.H:119-131);
        like the reference's ReactorBDF and ReactorRK64, the other integrators
        do not read them."""
        vals = [float(v) for v in typ_vals]
        if len(vals) != self.nspec + 1:
            raise ReactorError(-1, f"set_typ_vals_ode: expected {self.nspec + 1} values, got {len(vals)}")
        self.typ_vals = vals
        arr = (C.c_double * len(vals))(*vals)
        rc = library().pp_chem_set_typ_vals(self._h, arr, len(vals))
        if rc != 0:
            raise ReactorError(rc, library().pp_chem_last_error(self._h).decode())

    def set_eos_parm(self, eosparm):
        """ReactorBase::set_eos_parm (ReactorBase.H:86-97): a parameter block
        whose identifier is "Manifold" is rejected, as the reference aborts for
        these reactors; any other (the ideal-gas EOS takes none) is kept."""
        if getattr(eosparm, "identifier", None) == "Manifold":
            raise ReactorError(-1, "Manifold EOS only supported with ReactorRK64 and ReactorNull for now")
        self.eosparm = eosparm

    def _set_manifold(self, parm) -> None:
        m, keep = parm.c_struct()
        rc = library().pp_chem_set_manifold(self._h, C.byref(m))
        del keep
        if rc != 0:
            raise ReactorError(rc, library().pp_chem_last_error(self._h).decode())
        self.eosparm = parm

    def close(self):
        if self._h:
            library().pp_chem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise ReactorError(rc, library().pp_chem_last_error(self._h).decode())
        return rc

    def react(self, rY, rYsrc, T, rE, rEsrc, FC, dt_react: float, time: float = 0.0,
              mask=None, stream: Optional[int] = None):
        """Box form on device tensors shaped (ncomp, nz, ny, nx); returns (avg_steps, time)."""
        self._box_tensors(rY=rY, rYsrc=rYsrc, T=T, rE=rE, rEsrc=rEsrc, FC=FC)
        nz, ny, nx = T.shape[-3:]
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        views = [_array4_from_tensor(t) for t in (rY, rYsrc, T, rE, rEsrc)]
        fc = _array4_from_tensor(FC) if FC is not None else Array4()
        dt = C.c_double(dt_react)
        tm = C.c_double(time)
        rc = library().pp_chem_react_box(self._h, lo, hi, *[C.byref(v) for v in views], C.byref(fc),
                                         None, C.byref(dt), C.byref(tm), stream)
        return self._check(rc), tm.value

    def _box_tensors(self, **views):
        """Every box view is a float64 GPU tensor over the same (nz, ny, nx)
        box as T, with the species count of components for rY / rYsrc."""
        T = views["T"]
        _device_f64(T, "T")
        box = tuple(T.shape[-3:])
        for name, t in views.items():
            if t is None and name == "FC":
                continue
            _device_f64(t, name)
            if t.dim() not in (3, 4) or tuple(t.shape[-3:]) != box:
                raise ValueError(f"{name}: expected shape (ncomp, {box[0]}, {box[1]}, {box[2]}), got {tuple(t.shape)}")
            ncomp = 1 if t.dim() == 3 else t.shape[0]
            need = self.nspec if name in ("rY", "rYsrc") else 1
            if ncomp < need:
                raise ValueError(f"{name}: expected at least {need} components, got {ncomp}")

    def flatten(self, rY, rYsrc, T, rE, rEsrc, stream: Optional[int] = None):
        """Box -> flat (YCOrder) state of device tensors shaped (ncomp, nz, ny, nx)
        (ReactorBase::flatten); returns torch tensors (y_vect, src_vect,
        vect_energy, src_vect_energy) on the same device."""
        import torch
        self._box_tensors(rY=rY, rYsrc=rYsrc, T=T, rE=rE, rEsrc=rEsrc)
        nz, ny, nx = T.shape[-3:]
        ncells, ns = nx * ny * nz, self.nspec
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        views = [_array4_from_tensor(t) for t in (rY, rYsrc, T, rE, rEsrc)]
        opts = dict(dtype=torch.float64, device=T.device)
        y, ys = torch.empty(ncells * (ns + 1), **opts), torch.empty(ncells * ns, **opts)
        e, es = torch.empty(ncells, **opts), torch.empty(ncells, **opts)
        rc = library().pp_chem_flatten(self._h, lo, hi, *[C.byref(v) for v in views], y.data_ptr(),
                                       ys.data_ptr(), e.data_ptr(), es.data_ptr(), stream)
        self._check(rc)
        return y, ys, e, es

    def unflatten(self, rY, T, rE, rEsrc, FC, y_vect, vect_energy, FCunt, dt_react: float,
                  stream: Optional[int] = None):
        """Flat -> box state (ReactorBase::unflatten): rY, T, rE, FC updated in
        place; FCunt is an int64 device tensor of per-cell step counts."""
        import torch
        self._box_tensors(rY=rY, T=T, rE=rE, rEsrc=rEsrc, FC=FC)
        nz, ny, nx = T.shape[-3:]
        ncells = nx * ny * nz
        _device_f64(y_vect, "y_vect", ncells * (self.nspec + 1))
        _device_f64(vect_energy, "vect_energy", ncells)
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        views = [_array4_from_tensor(t) for t in (rY, T, rE, rEsrc, FC)]
        if not isinstance(FCunt, torch.Tensor) or FCunt.dtype != torch.int64 or FCunt.device.type != "cuda":
            raise ValueError("FCunt must be an int64 GPU tensor")
        if FCunt.numel() != ncells:
            raise ValueError(f"FCunt: expected {ncells} elements, got {FCunt.numel()}")
        rc = library().pp_chem_unflatten(self._h, lo, hi, *[C.byref(v) for v in views], y_vect.data_ptr(),
                                         vect_energy.data_ptr(), FCunt.data_ptr(), C.c_double(dt_react),
                                         stream)
        self._check(rc)

    def react_1d(self, y, ysrc, rX, rXsrc, dt_react: float, time: float = 0.0,
                 stream: Optional[int] = None):
        """1-D form: torch device tensors (in place) or numpy host arrays (copied)."""
        ncells = int(rX.shape[0])
        sizes = self._sizes_1d(ncells)
        if isinstance(y, np.ndarray):
            for a, (name, size) in zip((y, ysrc, rX, rXsrc), sizes):
                if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError(f"{name}: host arrays must be contiguous float64 numpy arrays")
                if a.size != size:
                    raise ValueError(f"{name}: expected {size} elements, got {a.size}")
            ptrs = [a.ctypes.data for a in (y, ysrc, rX, rXsrc)]
            on_dev = 0
        else:
            self._device_1d((y, ysrc, rX, rXsrc), sizes)
            ptrs = [a.data_ptr() for a in (y, ysrc, rX, rXsrc)]
            on_dev = 1
        dt = C.c_double(dt_react)
        tm = C.c_double(time)
        rc = library().pp_chem_react_1d(self._h, *ptrs, C.byref(dt), C.byref(tm), ncells, on_dev,
                                        stream)
        return self._check(rc), tm.value

    def react_1d_noalloc(self, y, ysrc, rX, rXsrc, dt_react: float, time: float = 0.0,
                         stream: Optional[int] = None):
        """The 1-D form on device tensors through pp_chem_react_1d_noalloc: no
        device allocation (a per-GPU static scratch array holds the queue,
        step counts and workspace; concurrent calls on one GPU race).
        Returns (avg steps, time)."""
        for a in (y, ysrc, rX, rXsrc):
            if isinstance(a, np.ndarray):
                raise ValueError("react_1d_noalloc takes contiguous device tensors")
        ncells = int(rX.shape[0])
        self._device_1d((y, ysrc, rX, rXsrc), self._sizes_1d(ncells))
        dt = C.c_double(dt_react)
        tm = C.c_double(time)
        rc = library().pp_chem_react_1d_noalloc(self._h, y.data_ptr(), ysrc.data_ptr(), rX.data_ptr(),
                                                rXsrc.data_ptr(), C.byref(dt), C.byref(tm), ncells, stream)
        return self._check(rc), tm.value

    def _sizes_1d(self, ncells: int):
        """(name, element count) of the YCOrder arrays y, ysrc, rX, rXsrc."""
        ns = self.nspec
        return (("y", ncells * (ns + 1)), ("ysrc", ncells * ns), ("rX", ncells), ("rXsrc", ncells))

    def _device_1d(self, arrays, sizes):
        for a, (name, size) in zip(arrays, sizes):
            _device_f64(a, name, size)
            if not a.is_contiguous():
                raise ValueError(f"{name}: device tensors must be contiguous")

    def react_1d_chained(self, y, ysrc, rX, rXsrc, dt_react: float, ncalls: int, time: float = 0.0,
                         stream: Optional[int] = None):
        """ncalls successive 1-D calls on device tensors, replayed from a HIP
        graph of one call; returns (avg steps of the last call, time).  Raises
        ReactorError (EINTEGRATE) when a cell failed in any of the calls."""
        for a in (y, ysrc, rX, rXsrc):
            if isinstance(a, np.ndarray):
                raise ValueError("chained calls take contiguous device tensors")
        self._device_1d((y, ysrc, rX, rXsrc), self._sizes_1d(int(rX.shape[0])))
        dt = C.c_double(dt_react)
        tm = C.c_double(time)
        rc = library().pp_chem_react_1d_chained(self._h, y.data_ptr(), ysrc.data_ptr(), rX.data_ptr(),
                                                rXsrc.data_ptr(), C.byref(dt), C.byref(tm),
                                                int(rX.shape[0]), int(ncalls), stream)
        return self._check(rc), tm.value

    def adjoint_1d(self, y, rX, nmacro: int, dt_react: float, species=("CO", "OH"),
                   stream: Optional[int] = None):
        """Fixed-step BDF over nmacro calls of dt_react (this handle's ode.bdf_*
        keys, exact linear solves) with its discrete adjoint.  y: (ncells, N+1)
        float64 device tensor, overwritten with the final state; rX: (ncells,)
        rho e / rho h.  Returns (out, grad): out (ncells, 3) = T_final, Y_a, Y_b;
        grad (ncells, NR, 3) = d out / d ln A_j, j in mechanism order."""
        import torch
        from .mechgen.__main__ import load
        ncells = int(rX.shape[0])
        self._device_1d((y, rX), (("y", ncells * (self.nspec + 1)), ("rX", ncells)))
        names = [sp.name for sp in load(self.mechanism.replace("_table", "")).species]
        sa, sb = (names.index(k) for k in species)
        nr = library().pp_chem_num_reactions(self._h)
        out = torch.empty((ncells, 3), dtype=torch.float64, device=y.device)
        grad = torch.empty((ncells, nr, 3), dtype=torch.float64, device=y.device)
        rc = library().pp_chem_adjoint_1d(self._h, y.data_ptr(), rX.data_ptr(), ncells, nmacro, dt_react, sa, sb,
                                          out.data_ptr(), grad.data_ptr(), stream)
        self._check(rc)
        return out, grad

    def production_rate(self, sc, T, stream: Optional[int] = None):
        """productionRate (Mechanisms/<mech>/mechanism.H) of the handle's
        mechanism on the GPU: sc (ncells, N) in mol/m^3 and T (ncells,) in K,
        float64 device tensors; returns wdot (ncells, N) in mol/m^3/s."""
        import torch
        ncells = int(T.shape[0])
        _device_f64(sc, "sc", ncells * self.nspec)
        _device_f64(T, "T", ncells)
        w = torch.empty((ncells, self.nspec), dtype=torch.float64, device=T.device)
        self._check(library().pp_chem_production_rate(self._h, sc.data_ptr(), T.data_ptr(), ncells,
                                                      w.data_ptr(), stream))
        return w

    def progress_rate(self, sc, T, stream: Optional[int] = None):
        """Net rates of progress q_f - q_r (comp_qfqr of Mechanisms/<mech>/mechanism.H)
        on the GPU; returns q (ncells, NR) in mol/m^3/s."""
        import torch
        ncells = int(T.shape[0])
        _device_f64(sc, "sc", ncells * self.nspec)
        _device_f64(T, "T", ncells)
        nr = library().pp_chem_num_reactions(self._h)
        q = torch.empty((ncells, nr), dtype=torch.float64, device=T.device)
        self._check(library().pp_chem_progress_rate(self._h, sc.data_ptr(), T.data_ptr(), ncells, q.data_ptr(),
                                                    stream))
        return q

    def ajacobian(self, sc, T, consP: int, stream: Optional[int] = None):
        """aJacobian (Mechanisms/<mech>/mechanism.H) on the GPU; returns J
        (ncells, N+1, N+1) with J[c, k, m] = d wdot_m / d c_k (column-major
        per cell, as the reference's J[k*(N+1)+m])."""
        import torch
        ncells = int(T.shape[0])
        n = self.nspec + 1
        _device_f64(sc, "sc", ncells * self.nspec)
        _device_f64(T, "T", ncells)
        J = torch.empty((ncells, n, n), dtype=torch.float64, device=T.device)
        self._check(library().pp_chem_ajacobian(self._h, sc.data_ptr(), T.data_ptr(), ncells, int(consP),
                                                J.data_ptr(), stream))
        return J

    def stats(self) -> dict:
        s = Stats()
        library().pp_chem_get_stats(self._h, C.byref(s))
        return s.as_dict()

    def workspace_rows(self) -> int:
        """Doubles per resident cell of the per-cell matrix workspace (0: none)."""
        return int(library().pp_chem_workspace_rows(self._h))

    def cell_steps(self, ncells: int) -> np.ndarray:
        out = np.zeros(ncells, np.int32)
        self._check(library().pp_chem_copy_cell_steps(self._h, out.ctypes.data, ncells))
        return out


class ReactorBDF(ReactorBase):
    identifier = "ReactorBDF"
    _mode = BDF


class ReactorRK64(ReactorBase):
    identifier = "ReactorRK64"
    _mode = RK64

    def set_eos_parm(self, eosparm):
        """ReactorRK64::set_eos_parm (ReactorRK64.H:123-128) keeps the EOS
        parameters for its RHS: with mechanism "Manifold" a
        pelephysics_amd.manifold.ManifoldParm (the table every lookup of the
        next calls interpolates in); the ideal-gas EOS takes none."""
        if getattr(eosparm, "identifier", None) == "Manifold":
            if self.mechanism != "Manifold":
                raise ReactorError(-1, "the Manifold EOS needs mechanism \"Manifold\" (Chemistry_Model=Null)")
            self._set_manifold(eosparm)
        else:
            self.eosparm = eosparm


class ReactorBDFAdaptive(ReactorBase):
    identifier = "ReactorBDFAdaptive"
    _mode = BDF_ADAPTIVE


class ReactorBDFAdaptiveMixed(ReactorBase):
    """Batched adaptive BDF: persistent work-queue kernel, FP32 Newton-matrix LU."""
    identifier = "ReactorBDFAdaptiveMixed"
    _mode = BDF_ADAPTIVE_MIXED


class ReactorBDFBatchedFP64(ReactorBase):
    """Batched adaptive BDF with an FP64 LU (isolates the FP32 LU's effect)."""
    identifier = "ReactorBDFBatchedFP64"
    _mode = BDF_BATCHED_FP64


class ReactorCvode(ReactorBase):
    """ReactorCvode's options and defaults (ReactorCvode.H:179-192: rtol 1e-6,
    atol 1e-10 x typical values, order <= 2, 3 Newton iterations, Jacobian
    age <= 100 steps, 10000 steps) on the FP64 adaptive BDF, one cell at a
    time as the reference's CPU path integrates; FC and the return value
    count RHS evaluations."""
    identifier = "ReactorCvode"
    _mode = CVODE


class ReactorArkode(ReactorBase):
    """ReactorArkode's explicit Runge-Kutta tables (ode.rk_method, default 40:
    ARKODE_ZONNEVELD_5_3_4) and step controllers (ode.rk_controller, default
    0: PID) with rtol 1e-6 / atol 1e-10 (x typical values), each cell with its
    own step size; FC and the return value count RHS evaluations."""
    identifier = "ReactorArkode"
    _mode = ARKODE


_REGISTRY = {c.identifier: c for c in (ReactorBDF, ReactorRK64, ReactorBDFAdaptive,
                                       ReactorBDFAdaptiveMixed, ReactorBDFBatchedFP64, ReactorCvode,
                                       ReactorArkode)}
