"""CPU checks of the drop-in boundary: the library loads, exports every entry point that
include/imogen_hip.h declares, and rejects bad configurations the way the reference
does (no GPU call is made)."""
import ctypes
import os
import re

import pytest

from imogen_amd import _lib, config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "imogen_hip.h")
EXT_HEADER = os.path.join(ROOT, "include", "imogen_hip_ext.h")


def declared_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(imh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["imh_create", "imh_destroy", "imh_upload_state", "imh_download_field",
                     "imh_fluid_sweep", "imh_magnet_sweep", "imh_time_update", "imh_flux",
                     "imh_iteration", "imh_run", "imh_batch_create", "imh_batch_step"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions() + declared_functions(EXT_HEADER)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function also has a ctypes signature in the binding
    assert set(declared_functions()) <= set(_lib.SIGNATURES)
    assert set(declared_functions(EXT_HEADER)) <= set(_lib.EXTRA)


def test_version_and_error_string():
    L = _lib.lib()
    assert L.imh_version().startswith(b"imogen-hip")
    assert isinstance(L.imh_last_error(), bytes)


def test_create_rejects_bad_configs_before_touching_the_device():
    L = _lib.lib()
    from imogen_amd.imogen import make_imh_config
    h = ctypes.c_void_p()
    # (10, 1, 8) squeezes to (10, 8): x's fade needs 20 cells along the first dimension
    rc = config.RunConfig(grid=(10, 1, 8), dgrid=(1, 1, 1),
                          bc=[["fade", "circ", "circ"], ["fade", "circ", "circ"]])
    assert L.imh_create(ctypes.byref(make_imh_config(rc)), ctypes.byref(h)) == -6
    rc = config.RunConfig(grid=(10, 10, 1), dgrid=(1, 1, 1),
                          bc=[["fade"] * 3, ["fade"] * 3])     # fade needs >= 20 cells
    assert L.imh_create(ctypes.byref(make_imh_config(rc)), ctypes.byref(h)) == -6
    assert b"fade" in L.imh_last_error()
    rc = config.RunConfig(grid=(10, 10, 1), dgrid=(1, 1, 1),
                          bc=[["wormhole", "circ", "circ"], ["const", "circ", "circ"]])
    assert L.imh_create(ctypes.byref(make_imh_config(rc)), ctypes.byref(h)) == -6
    assert b"3-D grid" in L.imh_last_error()   # wormhole_shift.m indexes a third dimension
    rc = config.RunConfig(grid=(10, 10, 4), dgrid=(1, 1, 1), magnet_active=True,
                          bc=[["circ", "circ", "wormhole"], ["circ", "circ", "wormhole"]])
    assert L.imh_create(ctypes.byref(make_imh_config(rc)), ctypes.byref(h)) == -6
    assert b"magnet" in L.imh_last_error()
    rc = config.RunConfig(grid=(10, 10, 1), dgrid=(1, 1, 1))
    c = make_imh_config(rc)
    c.limiter[1] = 7
    assert L.imh_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert L.imh_create(None, ctypes.byref(h)) == -1


def test_config_parsing_mirrors_initialize():
    with pytest.raises(ValueError, match="Unknown flux limiter type: minmod"):
        config.from_ini({"grid": (4, 4, 1), "fluxLimiter": "minmod"})
    rc = config.from_ini({"grid": (4, 4, 1), "fluxLimiter": config.FluxLimiterEnum.MINMOD,
                          "bcMode": {"x": ["trans", "const"], "y": "circ"},
                          "mode": {"magnet": True}})
    assert rc.limiter == ["minmond"] * 3
    assert rc.bc == [["trans", "circ", "circ"], ["const", "circ", "circ"]]
    assert rc.cfl == 0.35  # Initializer.m:143-150 default with the magnet on
    with pytest.raises(ValueError, match="Unknown BC mode"):
        config.from_ini({"grid": (4, 4, 1), "bcMode": "linear"})
    assert config.from_ini({"grid": (4, 4, 4), "bcMode": {"z": "wormhole"}}).bc[0][2] == "wormhole"
