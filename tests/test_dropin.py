"""The drop-in boundary on the CPU: the reference's C++-linkage entry points are
exported by libtsr.so under the reference's own (mangled) names, the reference's
unchanged driver compiles and links against it (alone, and with the reference's own
loaders and save_solution), and a caller including the reference's headers can include
the library's header too."""
import ctypes as C
import os
import subprocess

import pytest

from twoscalerich_amd import _ffi as F

from dropin_case import DRIVERS, ROOT

REF = "/root/reference"
HAVE_REF = os.path.isdir(os.path.join(REF, "src"))

# include/main.h:21-34 and include/exprem.h:27-30, as g++ mangles them
MANGLED = {
    "Gfunc_serial": "_Z12Gfunc_serialPdS_16user_data_struct",
    "Gfunc_parallel": "_Z14Gfunc_parallelPdS_16user_data_struct",
    "macro_mesh_properties": "_Z21macro_mesh_propertiesPKcR17macro_mesh_structPi",
    "micro_mesh_properties": "_Z21micro_mesh_propertiesPKcR17micro_mesh_struct",
    "read_effective_conductivity": "_Z27read_effective_conductivityPKcR11Keff_struct",
    "set_exprem_options": "_Z18set_exprem_optionsR21exprem_options_struct",
    "exprem": "_Z6expremPFiPdS_16user_data_structEiRddS_S_S_S0_21exprem_options_structR19exprem_stats_struct",
    "save_solution": "_Z13save_solutionNSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEEPdS5_i16user_data_struct",
    "weighted_norm": "_Z13weighted_normPdS_ddi",
    "phipade": "_Z7phipadePdS_dii",
    "check_step": "_Z10check_stepdd",
    "check_flag_Gfunc": "_Z16check_flag_Gfuncid",
}


def test_reference_cxx_symbols_are_exported():
    L = C.CDLL(F.LIB_PATH)
    missing = [n for n, m in MANGLED.items() if not hasattr(L, m)]
    assert not missing, missing
    # and the C-linkage twins of the helpers
    for n in ("phipade", "weighted_norm", "check_step", "check_flag_Gfunc", "Gfunc_hip", "exprem_c"):
        assert hasattr(L, n), n


def _cxx(args, cwd):
    r = subprocess.run(["g++", "-O1", "-w", "-fopenmp"] + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


LINK = ["-L" + os.path.join(ROOT, "twoscalerich_amd"), "-ltsr",
        "-Wl,-rpath," + os.path.join(ROOT, "twoscalerich_amd")]


@pytest.mark.skipif(not HAVE_REF, reason="reference tree absent")
def test_reference_driver_links_unchanged(tmp_path):
    """src/main.cpp, not one line changed, against libtsr.so alone (the library supplies
    exprem, both RHS names, the loaders, the options and save_solution), and together with
    the reference's own loaders (every reference TU but phipade.cpp, which needs Apple
    Accelerate): both link."""
    src = os.path.join(REF, "src")
    _cxx(["-I" + os.path.join(REF, "include"), os.path.join(src, "main.cpp"), "-o", "drv"] + LINK, tmp_path)
    io = [os.path.join(src, f + ".cpp") for f in ("macro_mesh_properties", "micro_mesh_properties",
                                                   "read_effective_conductivity", "set_exprem_options",
                                                   "save_solution")]
    _cxx(["-I" + os.path.join(REF, "include"), os.path.join(src, "main.cpp")] + io + ["-o", "drv_io"] + LINK, tmp_path)
    # it loads libtsr.so and starts: a missing options file ends it with main.cpp:40-43's 1
    r = subprocess.run([str(tmp_path / "drv"), "no_such_file.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Reading configuration file" in r.stdout


@pytest.mark.skipif(not HAVE_REF, reason="reference tree absent")
def test_reference_headers_and_library_header_together(tmp_path):
    """A caller that includes the reference's main.h (which defines the structs, without
    include guards) and then the library's header with TSR_USE_REFERENCE_HEADERS compiles
    and links, and can take the device RHS under either name."""
    (tmp_path / "caller.cpp").write_text(
        '#include "main.h"\n'
        '#define TSR_USE_REFERENCE_HEADERS\n'
        '#include "tsr/tsr.h"\n'
        'int main() {\n'
        '    int (*G)(double*, double*, user_data_struct) = Gfunc_hip;\n'
        '    int (*P)(double*, double*, user_data_struct) = Gfunc_parallel;\n'
        '    exprem_options_struct o; set_exprem_options(o);\n'
        '    return (G != nullptr && P != nullptr && o.MaxKrylov > 0) ? 0 : 1;\n'
        '}\n')
    _cxx(["-I" + os.path.join(REF, "include"), "-I" + os.path.join(ROOT, "include"), "caller.cpp", "-o", "caller"]
         + LINK, tmp_path)
    assert subprocess.run([str(tmp_path / "caller")], timeout=60).returncode == 0


@pytest.mark.skipif(not HAVE_REF, reason="reference tree absent")
def test_prebuilt_drivers_present():
    """build() leaves both drivers in oracle/_ref/ for the GPU tests (which cannot read
    the reference tree)."""
    for name, exe in DRIVERS.items():
        assert os.access(exe, os.X_OK), name
