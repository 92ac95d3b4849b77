"""CPU-only checks of the C-ABI library: it loads, exports every declared entry
point, and the ctypes mirrors of the public structs match the C layout."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sph.h")
LIB = os.path.join(ROOT, "jax-sph_amd", "jax_sph_amd", "libsph.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|size_t|const char \*)\s*\*?\s*(sph_\w+)\s*\(",
                                 text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("sph_ws_create", "sph_nl_build", "sph_nl_write_idx", "sph_forward",
                 "sph_step", "sph_kernel_eval", "sph_integrate", "sph_cell_list",
                 "sph_cell_list_work_bytes"):
        assert must in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsph.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsph.so not built")
def test_ctypes_struct_layout_matches_c():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "jax-sph_amd"))
    from jax_sph_amd import _lib
    src = r'''
    #include <stdio.h>
    #include <stddef.h>
    #include "sph.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu\n", sizeof(sph_nl_config), sizeof(sph_solver_params),
             sizeof(sph_state_in), sizeof(sph_state_out), sizeof(sph_bc_spec));
      printf("%zu %zu %zu\n", offsetof(sph_nl_config, ell_width),
             offsetof(sph_solver_params, tvf_stress), offsetof(sph_state_in, tag));
      printf("%zu %zu %zu %zu\n", sizeof(sph_ensemble_member),
             offsetof(sph_ensemble_member, r_out), offsetof(sph_ensemble_member, max_occupancy),
             offsetof(sph_ensemble_member, info));
      return 0;
    }
    '''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        r = subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                            "-I", "/opt/rocm/include", c, "-o", exe], capture_output=True,
                           text=True)
        if r.returncode != 0:
            pytest.skip("cannot compile the layout probe: " + r.stderr[-300:])
        out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    sizes = list(map(int, out))
    assert sizes[:5] == [ctypes.sizeof(_lib.NLConfig), ctypes.sizeof(_lib.SolverParams),
                         ctypes.sizeof(_lib.StateIn), ctypes.sizeof(_lib.StateOut),
                         ctypes.sizeof(_lib.BCSpec)]
    assert sizes[5:8] == [_lib.NLConfig.ell_width.offset, _lib.SolverParams.tvf_stress.offset,
                          _lib.StateIn.tag.offset]
    M = _lib.EnsembleMember
    assert sizes[8:] == [ctypes.sizeof(M), M.r_out.offset, M.max_occupancy.offset,
                         M.info.offset]


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "jax-sph_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
