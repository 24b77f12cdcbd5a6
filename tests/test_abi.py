"""The C-ABI library loads without a GPU and exports every function that
include/tinyfluids.h declares; argument checks answer before any device work."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tinyfluids.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(tf_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("tf_evolve_state", "tf_cfl_time_step", "tf_time_integration", "tf_conserved_state_from_primitive",
              "tf_primitive_state_from_conserved", "tf_evolve_state_along_axis", "tf_open_boundaries",
              "tf_batch_create", "tf_batch_step_graph"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from tinyfluids_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for every exported function
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_header_is_plain_c_and_links(tmp_path):
    """include/tinyfluids.h compiles as C99 with warnings as errors, a C
    program that takes the address of every declared entry point links
    against libtfhip.so, and it runs without a GPU for the calls that need
    none (version, error string, argument checks)."""
    import shutil
    import subprocess
    from tinyfluids_amd import _lib
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lib_path = _lib.load()._name
    names = declared_functions()
    src = tmp_path / "abi_check.c"
    src.write_text(
        '#include <stdio.h>\n#include "tinyfluids.h"\n'
        "typedef void (*fn_t)(void);\n"
        "static fn_t table[] = {" + ", ".join(f"(fn_t){n}" for n in names) + "};\n"
        "int main(void) {\n"
        "    if (tf_abi_version() != TF_ABI_VERSION) return 1;\n"
        "    if (tf_evolve_state(0, 0, 4, 4, 4, 0.1, 0, 1.4, 0, 0, 0, 0) != TF_ERR_ARG) return 2;\n"
        "    printf(\"%d %s\\n\", (int)(sizeof table / sizeof table[0]), tf_last_error());\n"
        "    return 0;\n}\n")
    exe = tmp_path / "abi_check"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{os.path.join(ROOT, 'include')}",
                    str(src), lib_path, f"-Wl,-rpath,{os.path.dirname(lib_path)}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == len(names)
    assert "null" in " ".join(out[1:])


def test_argument_errors_without_gpu():
    from tinyfluids_amd import _lib
    lib = _lib.load()
    assert lib.tf_abi_version() == 1
    assert lib.tf_evolve_state(None, None, 4, 4, 4, 0.1, None, 1.4, 0, None, None, None) == _lib.TF_ERR_ARG
    assert b"null" in lib.tf_last_error()
    assert lib.tf_evolve_state_along_axis(ctypes.c_void_p(16), ctypes.c_void_p(32), 4, 4, 4, 0.1,
                                          ctypes.c_void_p(48), 1.4, 7, None) == _lib.TF_ERR_ARG
    assert b"axis" in lib.tf_last_error()
    assert lib.tf_open_boundaries(ctypes.c_void_p(16), ctypes.c_void_p(16), 4, 4, 4, 1, None) == _lib.TF_ERR_ARG
    assert lib.tf_open_boundaries(ctypes.c_void_p(16), ctypes.c_void_p(32), 4, 4, 4, 4, None) == _lib.TF_ERR_ARG
    assert b"ghost" in lib.tf_last_error()
    assert lib.tf_open_boundaries(ctypes.c_void_p(16), ctypes.c_void_p(32), 4, 4, 4, -1, None) == _lib.TF_ERR_ARG
    assert lib.tf_evolve_state(ctypes.c_void_p(16), ctypes.c_void_p(32), 0, 4, 4, 0.1, ctypes.c_void_p(48), 1.4, 0,
                               None, None, None) == _lib.TF_ERR_ARG
    assert lib.tf_set_tiling(8, 8, 8, 3, 8) == _lib.TF_ERR_ARG
    assert lib.tf_workspace_bytes(8, 8, 8) >= 1024
    assert lib.tf_batch_workspace_bytes(4) > 0
    assert lib.tf_hbm_copy(ctypes.c_void_p(16), ctypes.c_void_p(4096), 24, None) == _lib.TF_ERR_ARG
    assert b"16" in lib.tf_last_error()
    assert lib.tf_hbm_copy(ctypes.c_void_p(16), ctypes.c_void_p(32), 64, None) == _lib.TF_ERR_ARG
    assert b"overlap" in lib.tf_last_error()
    with pytest.raises(ValueError):
        _lib.check(_lib.TF_ERR_ARG)


def test_batch_entry_points_check_arguments_without_gpu():
    """The batch plan and status entry points (synchronous and asynchronous
    forms) answer argument errors before any device work."""
    from tinyfluids_amd import _lib
    lib = _lib.load()
    handle = ctypes.c_void_p()
    for create in (lib.tf_batch_create, lib.tf_batch_create_async):
        assert create(ctypes.byref(handle), None, 0, ctypes.c_void_p(16), 1 << 20, 0, 0, None) == _lib.TF_ERR_ARG
        assert b"nsims" in lib.tf_last_error()
        assert create(None, None, 1, None, 0, 0, 0, None) == _lib.TF_ERR_ARG
        assert b"null" in lib.tf_last_error()
    assert lib.tf_batch_status_bytes(None) == 0
    assert lib.tf_batch_status_enqueue(None, ctypes.c_void_p(16), None) == _lib.TF_ERR_ARG
    assert b"null" in lib.tf_last_error()
    out = (_lib.SimStatus * 1)()
    assert lib.tf_batch_status_unpack(None, ctypes.c_void_p(16), out) == _lib.TF_ERR_ARG
    assert lib.tf_batch_status(None, out, None) == _lib.TF_ERR_ARG


def test_python_front_end_rejects_bad_states():
    import torch
    import tinyfluids_amd as tf
    with pytest.raises(TypeError):
        tf._evolve_state(torch.zeros(5, 4, 4, 4, dtype=torch.float64), 0.1, 0.01, 1.4)
    with pytest.raises(TypeError):
        tf._evolve_state(torch.zeros(4, 4, 4), 0.1, 0.01, 1.4)
    with pytest.raises(TypeError):  # no CPU path
        tf._evolve_state(torch.zeros(5, 4, 4, 4), 0.1, 0.01, 1.4)
    with pytest.raises(NotImplementedError):
        tf.time_integration(torch.zeros(5, 4, 4, 4), 0.1, 0.2, 1.4, False, True)
