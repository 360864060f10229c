"""CPU checks of the C-ABI boundary: the library loads and exports every
function include/pm355.h declares; without a GPU, context creation reports
PM_ERR_NO_DEVICE instead of crashing or exiting."""
import ctypes as C

import pm355


def test_library_exports_every_declared_symbol():
    lib = pm355.load_library()
    names = pm355.declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point also has a ctypes signature in the harness
    assert set(names) <= set(pm355._SIGS), set(names) - set(pm355._SIGS)


def test_version_and_no_device_status():
    import torch

    lib = pm355.load_library()
    assert b"gfx950" in lib.pm_version()
    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    rc = lib.pm_ctx_create(0, None, C.byref(h))
    assert rc == pm355.PM_ERR_NO_DEVICE and not h.value


def test_null_context_is_rejected():
    lib = pm355.load_library()
    assert lib.pm_ctx_sync(None) == pm355.PM_ERR_INVALID
    assert lib.pm_grid_destroy(None) == pm355.PM_OK
    assert lib.pm_last_error(None) == b"null context"
