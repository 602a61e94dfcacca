"""The C-ABI library loads and exports every symbol include/sslab_hip.h declares
(CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sslab_hip.h")
LIB = os.path.join(ROOT, "sslablu_amd", "libsslab_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sslab_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "sslab_rb_factor" in syms and "sslab_rb_solve" in syms
    assert len(syms) >= 12


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (load order: torch's HIP runtime first)
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.sslab_abi_version.restype = ctypes.c_int
    assert lib.sslab_abi_version() >= 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_header():
    from sslablu_amd import _lib
    assert set(declared_symbols()) <= set(_lib.EXPORTED)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_size_queries_and_argument_errors_without_gpu():
    """Size queries and argument validation run on the host."""
    from sslablu_amd import _lib
    f, w, i = _lib.sizes("rb", 31, 512)
    m2 = 512 * 512
    # levels 31 -> 16 -> 8 -> 4, then the 4-node coarsest level (4 * 512 = 2048 rows) as one
    # dense inverse: factor = sum(3*nO + 2*nE) m^2 + (4m)^2
    levels = [(31, 16, 15), (16, 8, 8), (8, 4, 4)]
    assert f == sum(3 * nO + 2 * nE for _, nE, nO in levels) * m2 + 16 * m2
    # m > 1024: the coarsest level is a single block; nb = 1: no level at all; nb >= 2
    # always eliminates once, even when the whole system would fit the dense limit
    assert _lib.sizes("rb", 3, 1500)[0] == (3 + 2 * 2 + 3 + 2) * 1500 ** 2 + 1500 ** 2
    assert _lib.sizes("rb", 1, 64)[0] == 64 * 64
    assert _lib.sizes("rb", 4, 16)[0] == (3 * 2 + 2 * 2) * 256 + (2 * 16) ** 2
    with pytest.raises(_lib.SslabError):
        _lib.sizes("rb", 0, 4)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_plan_mirror_matches_library():
    """blocktri.rb_levels (used for bench accounting) describes the same plan as the
    library: the factor size it implies equals sslab_rb_sizes for many (nb, m), and the
    backward-block count follows the same levels."""
    from sslablu_amd import _lib
    from sslablu_amd.blocktri import rb_backward_blocks, rb_levels
    for nb in (1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 31, 33, 64, 100):
        for m in (8, 100, 512, 700, 1024, 1500, 2048, 3000):
            L, n_top = rb_levels(nb, m)
            n, words = nb, 0
            for _ in range(L):
                words += 3 * (n // 2) + 2 * ((n + 1) // 2)
                n = (n + 1) // 2
            assert n == n_top
            assert _lib.sizes("rb", nb, m)[0] == words * m * m + (n_top * m) ** 2
            assert rb_backward_blocks(nb, m) <= 2 * sum(1 for _ in range(L)) * nb


def test_grouped_gemm_struct_layout_matches_header(tmp_path):
    """The ctypes mirrors of sslab_gemm_term / sslab_gemm_desc (sslablu_amd/hbs.py) have
    the C layout of include/sslab_hip.h (offsets and sizes from a gcc-compiled probe)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from sslablu_amd.hbs import _GDesc, _GTerm
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "sslab_hip.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu\\n", sizeof(sslab_gemm_term), sizeof(sslab_gemm_desc));\n'
        '  printf("%zu %zu %zu %zu %zu\\n", offsetof(sslab_gemm_term, K), offsetof(sslab_gemm_term, alpha),'
        " offsetof(sslab_gemm_desc, nterms), offsetof(sslab_gemm_desc, Cin), offsetof(sslab_gemm_desc, ldcin));\n"
        '  printf("%zu %zu %zu\\n", offsetof(sslab_gemm_desc, C), offsetof(sslab_gemm_desc, ldc),'
        " offsetof(sslab_gemm_desc, beta));\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]
    import ctypes as ct
    assert vals == [ct.sizeof(_GTerm), ct.sizeof(_GDesc), _GTerm.K.offset, _GTerm.alpha.offset,
                    _GDesc.nterms.offset, _GDesc.Cin.offset, _GDesc.ldcin.offset,
                    _GDesc.C.offset, _GDesc.ldc.offset, _GDesc.beta.offset]
