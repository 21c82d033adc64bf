"""The C-ABI libraries load without a GPU and export every function their headers
declare (include/fjmi.h, include/fjmi_gen.h, include/cluster.h); the PseudoJet
record keeps the reference's 40-byte layout (PseudoJet.h:4-11 of the reference).
No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "fastjet_gpu_amd", "lib")


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fjmi_\w+)\s*\(", text, flags=re.M)))


def load(name):
    path = os.path.join(LIB, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    return ctypes.CDLL(path)


def test_header_parse_finds_every_entry_point():
    names = declared("fjmi.h")
    for must in ("fjmi_create", "fjmi_cluster_event_device", "fjmi_cluster_pseudojets", "fjmi_batch_create",
                 "fjmi_batch_launch", "fjmi_batch_stats", "fjmi_cluster_batch", "fjmi_server_create",
                 "fjmi_server_submit", "fjmi_server_poll", "fjmi_server_release", "fjmi_server_destroy",
                 "fjmi_stream_check", "fjmi_cluster_event_device_noalloc", "fjmi_inclusive_jets",
                 "fjmi_inclusive_pseudojets", "fjmi_history_view", "fjmi_exclusive_jets", "fjmi_cluster_event_def",
                 "fjmi_cluster_event_device_def", "fjmi_batch_create_def", "fjmi_exclusive_jets_def",
                 "fjmi_cluster_batch_def"):
        assert must in names
    assert len(names) == 36


def test_scheme_numbers_agree_with_the_kernels_header(tmp_path):
    """The public FJMI_SCHEME_* numbers and the kernels' (fjmi_math.h) are the same
    macros: a translation unit that includes both compiles without a redefinition
    warning, and the values are FastJet's RecombinationScheme order."""
    src = tmp_path / "schemes.c"
    src.write_text('#include "fjmi.h"\n#include "fjmi_math.h"\n#include <stdio.h>\n'
                   'int main(void) { printf("%d %d %d %d %d %d %d %d %d %d\\n", FJMI_SCHEME_E, FJMI_SCHEME_PT, '
                   'FJMI_SCHEME_PT2, FJMI_SCHEME_ET, FJMI_SCHEME_ET2, FJMI_SCHEME_BIPT, FJMI_SCHEME_BIPT2, '
                   'FJMI_SCHEME_WTA_PT, FJMI_SCHEME_WTA_MODP, FJMI_NUM_SCHEMES); '
                   'return (int)sizeof(fjmi_jetdef) - 16; }\n')
    exe = tmp_path / "schemes"
    subprocess.run(["gcc", "-std=c11", "-Werror", f"-I{INC}", f"-I{os.path.join(ROOT, 'fastjet_gpu_amd', 'csrc')}",
                    str(src), "-o", str(exe), "-lm"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == [str(k) for k in range(10)]


@pytest.mark.parametrize("header,lib", [("fjmi.h", "libfjmi.so"), ("fjmi_gen.h", "libfjmi_gen.so")])
def test_exports(header, lib):
    so = load(lib)
    missing = [n for n in declared(header) if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_dropin_cluster_symbol():
    """cluster(PseudoJet*, int, Algorithm, double) with C++ linkage, as cluster.h:6 declares it."""
    path = os.path.join(LIB, "libfjmi_cluster.so")
    if not os.path.exists(path):
        pytest.skip("libfjmi_cluster.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    assert "_Z7clusterP9PseudoJeti9Algorithmd" in out
    assert "_Z14inclusive_jetsPK9PseudoJetidPS_Pi" in out  # include/fjmi_dropin.h


def test_calls_without_gpu_are_errors_not_crashes():
    so = load("libfjmi.so")
    so.fjmi_version.restype = ctypes.c_int
    assert so.fjmi_version() >= 100
    so.fjmi_last_error.restype = ctypes.c_char_p
    so.fjmi_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    ctx = ctypes.c_void_p()
    rc = so.fjmi_create(-1, ctypes.byref(ctx))
    assert rc in (-1, -2) and so.fjmi_last_error()


def test_pseudojet_layout():
    """Compile a tiny C++ probe against include/PseudoJet.h: 40 bytes, fields at the reference's offsets."""
    src = r'''
#include <cstddef>
#include <cstdio>
#include "PseudoJet.h"
#include "cluster.h"
int main() {
  std::printf("%zu %zu %zu %zu %zu %zu %zu %d %d %d\n", sizeof(PseudoJet), offsetof(PseudoJet, index),
              offsetof(PseudoJet, isJet), offsetof(PseudoJet, px), offsetof(PseudoJet, py),
              offsetof(PseudoJet, pz), offsetof(PseudoJet, E), (int)Algorithm::Kt,
              (int)Algorithm::CambridgeAachen, (int)Algorithm::AntiKt);
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.cpp")
        open(c, "w").write(src)
        exe = os.path.join(d, "p")
        subprocess.run(["g++", "-std=c++17", f"-I{INC}", c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["40", "0", "4", "8", "16", "24", "32", "1", "0", "-1"]
