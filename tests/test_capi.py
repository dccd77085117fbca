"""The C-ABI library loads, exports every symbol declared in include/*.h, and
its host-side logic (config defaults, argument validation) works without a GPU."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names |= set(re.findall(r"^\w[\w\s\*]*?\b(pp_chem_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    from pelephysics_amd.reactor import library
    L = library()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n


def test_config_defaults_match_reference():
    from pelephysics_amd.reactor import library, Config
    L = library()
    cfg = Config()
    assert L.pp_chem_config_init(C.byref(cfg), b"drm19", 0, 2) == 0
    # ReactorBDF.H:118-128, ReactorRK64.H:131-134
    assert (cfg.bdf_nsubsteps, cfg.bdf_nonlinear_iters, cfg.bdf_gmres_restarts) == (1, 1, 10)
    assert (cfg.bdf_gmres_kspiters, cfg.bdf_gmres_precond, cfg.bdf_scheme) == (22, 2, 3)
    assert (cfg.bdf_nonlinear_tol, cfg.bdf_gmres_tol) == (1e-6, 1e-10)
    assert (cfg.rk64_atol, cfg.rk64_nsubsteps_guess, cfg.rk64_nsubsteps_min,
            cfg.rk64_nsubsteps_max) == (1e-10, 10, 5, 500)
    assert L.pp_chem_config_init(C.byref(cfg), b"nope", 0, 2) == -2
    # ReactorArkode.H:132-136 (ARKODE_ZONNEVELD_5_3_4, PID), ARKODE's 500 steps
    assert L.pp_chem_config_init(C.byref(cfg), b"drm19", 6, 2) == 0
    assert (cfg.rtol, cfg.atol, cfg.max_steps, cfg.erk_method, cfg.erk_controller) == (1e-6, 1e-10, 500, 40, 0)
    # the Manifold EOS: MANIFOLD_DIM 1 (Manifold.H:13-15)
    assert L.pp_chem_config_init(C.byref(cfg), b"Manifold", 1, 2) == 0
    assert cfg.manifold_dim == 1 and cfg.mechanism == b"Manifold"


def test_invalid_configs_rejected_before_device_use():
    from pelephysics_amd.reactor import library, Config
    L = library()
    h = C.c_void_p()
    for mutate in (lambda c: setattr(c, "reactor_type", 3),
                   lambda c: setattr(c, "bdf_scheme", 4),
                   lambda c: setattr(c, "bdf_gmres_kspiters", 99),
                   lambda c: setattr(c, "mode", 7),
                   lambda c: setattr(c, "block_size", 96)):
        cfg = Config()
        L.pp_chem_config_init(C.byref(cfg), b"LiDryer", 0, 2)
        mutate(cfg)
        assert L.pp_chem_create(C.byref(cfg), C.byref(h)) == -1
    # ReactorArkode: the tables not provided, an unknown controller, the
    # ReactorArkode keys on another integrator; the Manifold EOS on anything
    # but ReactorRK64 or with a dimension outside 1..5
    for mech, mode, mutate in (("LiDryer", 6, lambda c: setattr(c, "erk_method", 43)),
                               ("LiDryer", 6, lambda c: setattr(c, "erk_method", 54)),
                               ("LiDryer", 6, lambda c: setattr(c, "erk_controller", 4)),
                               ("LiDryer", 0, lambda c: setattr(c, "erk_method", 40)),
                               ("Manifold", 0, lambda c: None),
                               ("Manifold", 6, lambda c: None),
                               ("Manifold", 1, lambda c: setattr(c, "manifold_dim", 6))):
        cfg = Config()
        assert L.pp_chem_config_init(C.byref(cfg), mech.encode(), mode, 2) == 0
        mutate(cfg)
        assert L.pp_chem_create(C.byref(cfg), C.byref(h)) == -1, (mech, mode)


def test_mechanism_list_and_factory_names():
    from pelephysics_amd.reactor import mechanisms, ReactorBase, ReactorError
    assert set(mechanisms()) >= {"LiDryer", "drm19"}
    with pytest.raises(ReactorError):
        ReactorBase.create("ReactorNull")  # not provided (Factory.H:135-147 aborts likewise)
    assert ReactorBase.create("ReactorBDF").identifier == "ReactorBDF"
    assert ReactorBase.create("ReactorCvode").identifier == "ReactorCvode"
    assert ReactorBase.create("ReactorArkode").identifier == "ReactorArkode"
