"""The C-ABI library loads and exports every function declared in include/*.h,
and the Python mirror's LibrpaOptions layout matches the C header (CPU only)."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared():
    names = set()
    for f in os.listdir(INC):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(INC, f)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(librpa_\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared(lib):
    import librpa_amd
    out = subprocess.run(["nm", "-D", "--defined-only", librpa_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(_declared() - exported)
    assert not missing, missing
    assert len(_declared()) >= 55


def test_options_layout_matches_header(lib):
    import ctypes
    import librpa_amd
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "librpa.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(LibrpaOptions), offsetof(LibrpaOptions, nfreq),
         offsetof(LibrpaOptions, libri_chi0_threshold_C), offsetof(LibrpaOptions, option_output_Wc_Rf_mat));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", INC, c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    O = librpa_amd.Options
    assert vals == [ctypes.sizeof(O), O.nfreq.offset, O.libri_chi0_threshold_C.offset,
                    O.option_output_Wc_Rf_mat.offset]


def test_c_program_links_and_reports_errors(lib):
    # a C caller links against the library; without a GPU the Ec call records an error
    import librpa_amd
    src = r'''
#include <stdio.h>
#include "librpa.h"
int main(void) {
  librpa_amd_set_error_mode(1);
  LibrpaHandler *h = librpa_create_handler(MPI_COMM_WORLD);
  LibrpaOptions o; librpa_init_options(&o);
  double re[1], im[1];
  double ec = librpa_get_rpa_correlation_energy(h, &o, 1, re, im);
  char buf[512];
  int err = librpa_amd_get_last_error(buf, 512);
  printf("%d %d\n", err, ec != ec);
  librpa_destroy_handler(h);
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        libdir = os.path.dirname(librpa_amd.LIB_PATH)
        subprocess.run(["gcc", "-I", INC, c, "-o", exe, "-L", libdir, "-l:librpa_mi355x.so",
                        "-Wl,-rpath," + libdir], check=True)
        out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    assert out == ["1", "1"]


def test_precision_selection_values(lib):
    # chi0 precision 0..4 (3, 4 = mixed tau precision with FP32 / BF16x3 at the low-weight
    # points), Ec precision 0..2; anything else is an error
    import librpa_amd as L
    h = L.Handler()
    for chi0 in (L.FP64, L.FP32, L.BF16X3, L.MIXED, L.MIXED_BF16):
        for ec in (L.FP64, L.FP32, L.BF16X3):
            h.set_precision(chi0, ec)
    for chi0, ec in ((5, 0), (-1, 0), (0, 3), (3, 3)):
        with pytest.raises(L.LibrpaError, match="unknown precision"):
            h.set_precision(chi0, ec)
