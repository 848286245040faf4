"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, the Python mirror's struct layouts match the C header, and with no
GPU it fails loudly (no CPU fallback)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import sph2d
from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "sph2d.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sph2d_\w+)\s*\(", text, re.M)))


def test_header_and_python_mirror_list_the_same_entry_points():
    assert header_functions() == sorted(sph2d.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(sph2d.LIB_PATH):
        pytest.skip("libsph2d.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", sph2d.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sph2d_\w+)", out))
    missing = set(header_functions()) - exported
    assert not missing, missing
    lib = ctypes.CDLL(sph2d.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name)


def _c_layout(struct, fields):
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({struct}));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c11", "-o", exe, c], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("py,cname", [(sph2d.Params, "sph2d_params"), (sph2d.Diag, "sph2d_diag"),
                                     (sph2d.Scales, "sph2d_scales")])
def test_struct_layout_matches_header(py, cname):
    names = [f[0] for f in py._fields_]
    size, offs = _c_layout(cname, names)
    assert ctypes.sizeof(py) == size
    assert [getattr(py, n).offset for n in names] == offs


def test_no_gpu_fails_loudly():
    if gpu_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(sph2d.LIB_PATH):
        pytest.skip("libsph2d.so not built")
    p = sph2d.default_params()
    assert p.struct_size == ctypes.sizeof(sph2d.Params)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code in (3,)
    with pytest.raises(sph2d.SphError):
        sph2d.status_equation([1000.0])


def test_bad_parameters_are_rejected_before_any_device_call():
    p = sph2d.default_params(integrator=7)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 5
    p = sph2d.default_params(dt=-1.0)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 1
    p = sph2d.default_params()
    p.struct_size = 8
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 1


def test_tgv_params_scale_with_lattice():
    p = sph2d.tgv_params(2048)
    assert p.nx == p.ny == 2048
    assert p.dx == 1.0 / 2048 and p.h == 1.2 * p.dx and p.max_shift == 0.25 * p.dx
    assert p.dt == pytest.approx(1e-4 * 100 / 2048)


def test_fortran_driver_links_and_fails_loudly_without_gpu():
    exe = os.path.join(os.path.dirname(sph2d.LIB_PATH), "fortran", "tgv_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (flang absent)")
    if gpu_available():
        pytest.skip("a GPU is present")
    out = subprocess.run([exe, "16", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "runs on the GPU only" in out.stdout
