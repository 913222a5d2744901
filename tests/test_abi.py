"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gfx|gfo)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("include/gfx.h", "gf-only-sem_amd/libgfx.so"),
                                        ("include/gfx_io.h", "gf-only-sem_amd/libgfx.so"),
                                        ("include/gfx_surrogate.h", "gf-only-sem_amd/libgfx.so"),
                                        ("oracle/gfx_oracle.h", "oracle/liboracle.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) > 5
    so = C.CDLL(os.path.join(ROOT, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_crash():
    import pygfx
    if pygfx.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = pygfx.build_box(nx=2, ny=2, nz=2, lx=2e3, ly=2e3, lz=2e3, pml_xmin=0, pml_xmax=0,
                        pml_ymin=0, pml_ymax=0, pml_zmax=0)
    with pytest.raises(pygfx.GfxError) as ei:
        pygfx.Solver(m)
    assert ei.value.status == 2
    import numpy as np
    r = np.zeros(125 * 3)
    with pytest.raises(pygfx.GfxError):
        pygfx.element_residual(m.dxi_dx[:1125], m.jacobian[:125], m.lambda_[:125], m.mu_[:125],
                               np.zeros(25), np.zeros(5), 5, np.zeros(375), r)


def test_bad_arguments_rejected_before_device_use():
    import numpy as np
    import pygfx
    r = np.zeros(3)
    with pytest.raises(pygfx.GfxError) as ei:
        pygfx.element_residual(np.zeros(9), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1),
                               np.zeros(1), 1, np.zeros(3), r)
    assert ei.value.status == 1


def test_cpp_glue_header_compiles_and_links(tmp_path):
    """include/gfx_backend_hip.hpp is what a C++ caller includes; it must build
    with the host compiler alone and link against libgfx.so."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "gfx_backend_hip.hpp"\n'
                   'int main() { return gfx_hip::run_forward("q") == 1 ? 0 : 1; }\n')
    exe = tmp_path / "use"
    pkg = os.path.join(ROOT, "gf-only-sem_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe), "-L", pkg, "-lgfx", "-Wl,-rpath," + pkg], check=True)
    # a bad direction is rejected before any device or file access
    assert subprocess.run([str(exe)]).returncode == 0
