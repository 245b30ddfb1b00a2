"""The binding links into a ct2 build unchanged: a translation unit holding
ct2's own declarations (kick.hpp, drift.hpp, fft.hpp, particle.hpp) and a
reference to each links against ct2_gpu_binding.o, and the binding's object
defines the mangled names ct2's callers use (fft.hpp:19-20,28-29 take
non-const fft_float*).  Compiles and links only; nothing runs (no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINDING = os.path.join(ROOT, "ct2_amd", "compat", "ct2_gpu_binding.cpp")
DECLS = os.path.join(ROOT, "tests", "ct2_decls_link.cpp")


@pytest.fixture(scope="module")
def objs(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("link")
    ob, od = str(d / "binding.o"), str(d / "decls.o")
    for src, obj, inc in ((BINDING, ob, ["-I" + os.path.join(ROOT, "include"),
                                         "-I" + os.path.join(ROOT, "ct2_amd", "compat")]),
                          (DECLS, od, [])):
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-c", src, "-o", obj, *inc],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout
    return d, ob, od


def test_binding_defines_ct2_mangled_names(objs):
    _, ob, _ = objs
    r = subprocess.run(["nm", "--defined-only", ob], stdout=subprocess.PIPE, text=True, check=True)
    defined = {line.split()[-1] for line in r.stdout.splitlines() if line.strip()}
    for sym in ("_Z10fft_r2c_3dPdS_i", "_Z10fft_c2r_3dPdS_i", "_Z13transpose_r2cPdS_i", "_Z13transpose_c2rPdS_i",
                "_Z12transpose_yzPdi", "_Z8scramblePdiii", "_Z8fft_realPdii", "_Z4kickRK11kick_params",
                "_Z5driftff", "_Z17particle_allocatem", "_Z14particle_countv", "_Z14particles_initv",
                "_Z16particle_to_gridv", "_Z17particle_cell_dimv", "_Z17particle_get_celliii",
                "_Z23particle_cell_neighborsiii"):
        assert sym in defined, sym
    assert "_Z10fft_r2c_3dPKdPdi" not in defined


def test_ct2_declarations_link_against_the_binding(objs):
    d, ob, od = objs
    lib_dir = os.path.join(ROOT, "ct2_amd")
    assert os.path.exists(os.path.join(lib_dir, "libct2hip.so")), "build libct2hip.so first"
    exe = str(d / "ct2_link")
    r = subprocess.run(["g++", od, ob, f"-L{lib_dir}", "-lct2hip", "-o", exe],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
