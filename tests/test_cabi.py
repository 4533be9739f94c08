"""The C ABI library loads and exports every function include/mcpet.h declares; the struct
layouts seen from Python match the header (no device calls)."""
import ctypes
import os
import re
import subprocess

import mcpet
from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mcpet.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcpet_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", mcpet.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mcpet_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = mcpet.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_struct_layouts():
    assert ctypes.sizeof(mcpet.PsfRecord) == 44
    assert ctypes.sizeof(mcpet.ComptonTable) == 5460
    assert ctypes.sizeof(mcpet.RayleighTable) == 4 * 4 * 128 * 15 + 4 * 25005 * 15 + 2 * 128 * 15
    assert mcpet.version().startswith("mcpet-mi355x")


def test_struct_sizes_match_the_compiled_library():
    import ctypes as C
    mine = [C.sizeof(t) for t in (mcpet.Config, mcpet.Phantom, mcpet.Tables, mcpet.Tallies, mcpet.RunParams,
                                  mcpet.Stats, mcpet.PsfRecord, mcpet.Single, mcpet.ServerStats)]
    assert mcpet.struct_sizes() == mine
    assert mine[-2] == 32 and mcpet.SINGLE_DTYPE.itemsize == 32


def test_create_rejects_bad_arguments_without_touching_the_device():
    cfg = mcpet.Config()
    ph = mcpet.Phantom()
    tab = mcpet.Tables()
    ctx = ctypes.c_void_p()
    st = mcpet.lib().mcpet_create(ctypes.byref(cfg), ctypes.byref(ph), ctypes.byref(tab), ctypes.byref(ctx))
    assert st == 1 and not ctx.value
    assert b"null argument" in mcpet.lib().mcpet_last_error(None)


def test_surrogate_and_rebinning_entry_points_reject_bad_arguments_without_the_device():
    L = mcpet.lib()
    s = ctypes.c_void_p()
    for dims in ((30, 56, 56), (32, 0, 56), (2, 4, 4)):        # not multiples of 4 / empty / too small
        assert L.mcpet_surrogate_create(*dims, 2, 32, 1, ctypes.byref(s)) == 1 and not s.value
        assert b"multiples of 4" in L.mcpet_surrogate_last_error(None)
    assert L.mcpet_surrogate_create(32, 56, 56, 0, 32, 1, ctypes.byref(s)) == 1          # no input channel
    assert L.mcpet_surrogate_create(32, 56, 56, 2, 32, 1, None) == 1                      # no output pointer
    out = (ctypes.c_float * 8)()
    assert L.mcpet_rebin_sinogram(None, 1, 2, 2, 2, out) == 1
    assert L.mcpet_surrogate_param_count(None) == 0
    # the plane map needs a config whose sizes were derived
    cfg = mcpet.Config()
    planes = (ctypes.c_int32 * 4)()
    assert L.mcpet_sinogram_planes(ctypes.byref(cfg), planes) == 1
    assert L.mcpet_server_fetch_rebinned(None, 1, None, None, None, None, None) == 1
    assert L.mcpet_project_maps(None, 4, 4, 4, None, None) == 1                       # no context
    assert L.mcpet_server_submit_estimate(None, None, None, 0, None, None) == 1        # no network
    assert L.mcpet_server_fetch_estimate(None, 1, None, None, None) == 1
    assert L.mcpet_surrogate_forward_ms(None, None) == 1
    assert L.mcpet_batch_project_maps(None, 0, 4, 4, 4, None, None) == 1
