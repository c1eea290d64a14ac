"""C ABI checks that need no GPU: the library loads, exports every function
include/lbm_c_api.h declares, and the ctypes struct matches the C layout."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(lbm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("lbm_step", "lbm_moments", "lbm_equilibrium", "lbm_convert", "lbm_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from maddening_amd.build import library_path
    path = library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run `python -m maddening_amd.build`)")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    from maddening_amd import _lib
    assert set(_lib.exported_symbols()) <= set(declared_functions())


def test_version_and_error_string_without_gpu():
    from maddening_amd import _lib
    L = _lib.lib()
    assert L.lbm_version() == 8
    assert isinstance(L.lbm_last_error(), bytes)


def test_host_validation_errors_without_gpu():
    """Argument checks run before any HIP call."""
    from maddening_amd import _lib
    L = _lib.lib()
    d = _lib.LbmDesc()
    d.q, d.storage, d.tau = 19, 0, 0.5
    d.dims[0] = d.dims[1] = d.dims[2] = 8
    rc = L.lbm_step(None, ctypes.byref(d), None, None, None, None, None, None, None, None,
                    None, 1)
    assert rc == -1 and b"tau must be > 0.5" in L.lbm_last_error()
    d.tau = 0.8
    d.q = 15
    assert L.lbm_step(None, ctypes.byref(d), None, None, None, None, None, None, None, None,
                      None, 1) == -1
    assert b"q must be" in L.lbm_last_error()
    d.q = 19
    # exchanged-slot cotangents exist only for the reference wall rule
    d.bb_mode, d.flags = 2, _lib.LBM_FLAG_ADJ_IN_SWAPPED
    assert L.lbm_adjoint_step(None, ctypes.byref(d), None, 1, None, None, None, None, None, None,
                              None, None, None) == -2
    assert b"reference wall rule" in L.lbm_last_error()
    d.bb_mode, d.flags = 0, 0
    assert L.lbm_workspace_size(ctypes.byref(d), 10) == 8 ** 3 * 19 * 4
    assert L.lbm_state_bytes(ctypes.byref(d)) == 8 ** 3 * 19 * 4
    # halfway bounce-back adds two 4-byte link words per cell after the
    # state, at a 256-byte aligned offset
    d.bb_mode = 2
    assert L.lbm_workspace_size(ctypes.byref(d), 10) == 8 ** 3 * 19 * 4 + 8 ** 3 * 8
    d.storage, d.dims[0], d.dims[1], d.dims[2] = 2, 3, 5, 7
    assert L.lbm_workspace_size(ctypes.byref(d), 10) == 4096 + 105 * 8
    assert L.lbm_links_size(ctypes.byref(d)) == 105 * 8
    # link words are for halfway bounce-back only
    d.bb_mode = 1
    assert L.lbm_pull_phase(None, ctypes.byref(d), 1, ctypes.c_void_p(4096), ctypes.c_void_p(8192),
                            ctypes.c_void_p(4096), ctypes.c_void_p(4096), None, None, None, None,
                            None) == -1
    assert b"LBM_BB_HALFWAY only" in L.lbm_last_error()
    assert L.lbm_links_build(None, ctypes.byref(d), None, None) == -1


def test_batch_item_validation_without_gpu():
    """Batch items are checked before the plan is uploaded: a plane pitch
    below the cell count is refused."""
    from maddening_amd import _lib
    L = _lib.lib()
    items = (_lib.LbmBatchItem * 1)()
    it = items[0]
    it.q, it.bb_mode, it.tau = 19, 0, 0.8
    it.dims[0], it.dims[1], it.dims[2] = 5, 6, 7
    it.f[0], it.f[1] = 4096, 8192          # never dereferenced: validation fails first
    it.pitch = 5 * 6 * 7 - 1
    handle = ctypes.c_void_p()
    rc = L.lbm_batch_plan_create(None, items, 1, 1, ctypes.c_void_p(4096), ctypes.byref(handle),
                                 None)
    assert rc == -1 and b"pitch is below the cell count" in L.lbm_last_error()


def test_desc_layout_matches_c(tmp_path):
    from maddening_amd import _lib
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "lbm_c_api.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(lbm_desc),"
        " offsetof(lbm_desc, dims), offsetof(lbm_desc, tau), offsetof(lbm_desc, force),"
        " offsetof(lbm_desc, lid_velocity), offsetof(lbm_desc, block_x),"
        " offsetof(lbm_desc, flags), offsetof(lbm_desc, scheme));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    D = _lib.LbmDesc
    assert vals == [ctypes.sizeof(D), D.dims.offset, D.tau.offset, D.force.offset,
                    D.lid_velocity.offset, D.block_x.offset, D.flags.offset, D.scheme.offset]


def _c_layout(tmp_path, struct, fields):
    src = tmp_path / f"{struct}.c"
    fmt = " ".join(["%zu"] * (len(fields) + 1))
    args = ", ".join([f"sizeof({struct})"] + [f"offsetof({struct}, {f})" for f in fields])
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lbm_c_api.h"\n'
                   f'int main(void){{printf("{fmt}\\n", {args});return 0;}}\n')
    exe = tmp_path / struct
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    return [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]


@pytest.mark.parametrize("struct,cls", [("lbm_desc", "LbmDesc"), ("lbm_pipe_desc", "LbmPipeDesc"),
                                        ("lbm_batch_item", "LbmBatchItem")])
def test_every_struct_layout_matches_c(tmp_path, struct, cls):
    """Size and every field offset of each ABI struct, from gcc, against the
    ctypes mirror in maddening_amd/_lib.py."""
    from maddening_amd import _lib
    C = getattr(_lib, cls)
    names = [f[0] for f in C._fields_]
    vals = _c_layout(tmp_path, struct, names)
    assert vals == [ctypes.sizeof(C)] + [getattr(C, n).offset for n in names]
