"""The tiled kernel's source run on the CPU: one emulated 64-lane wavefront
(tests/emu/) under AddressSanitizer, every result compared bit-exactly with the
oracle.  Covers the kernel's bookkeeping (tile segments, labels, orphan rescans,
tile minima) without a GPU; the GPU tests run the same source on MI355X."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

EMU = os.path.join(ROOT, "tests", "emu")
LIB = os.path.join(ROOT, "fastjet_gpu_amd", "lib")
ORACLE = os.path.join(ROOT, "oracle", "_build")


def _build_emu(tmp_path_factory, name, defines=()):
    if not (os.path.exists(os.path.join(LIB, "libfjmi_gen.so")) and os.path.exists(os.path.join(ORACLE, "libfj_oracle.so"))):
        pytest.skip("generator or oracle library not built")
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("emu") / name)
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-ffp-contract=off",
           *defines,
           f"-I{EMU}", f"-I{os.path.join(ROOT, 'fastjet_gpu_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(ROOT, 'oracle')}", "-x", "c++", os.path.join(EMU, "tiled_emu_main.cpp"),
           os.path.join(EMU, "wave_fibers.cpp"), "-x", "none", f"-L{LIB}", "-lfjmi_gen", f"-L{ORACLE}", "-lfj_oracle",
           f"-Wl,-rpath,{LIB}", f"-Wl,-rpath,{ORACLE}", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def emu_bin(tmp_path_factory):
    return _build_emu(tmp_path_factory, "tiled_emu")


# The gather counter restarts every 40 gathers (instead of near 2^32), so every
# event restarts it many times, also between a step's gather and its orphan rescans.
STAMP_ENV = {"FJMI_EMU_STAMP_LIMIT": "40"}


# No segment slack: insertions relocate segments and a pool without spare room
# runs out, so the kernel's segment rebuild runs.
TIGHT_ENV = {"FJMI_EMU_SLACK_SHIFT": "30", "FJMI_EMU_SLACK_ADD": "0"}


def run(emu_bin, *args, env=None):
    r = subprocess.run([emu_bin, *map(str, args)], capture_output=True, text=True, timeout=600,
                       env=None if env is None else {**os.environ, **env})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("workload,e0,e1,variant,p,R", [
    (1, 0, 1, 0, -1, 0.4), (1, 0, 1, 0, 0, 0.4), (1, 0, 1, 0, 1, 0.4),
    (2, 0, 1, 0, 1, 1.0), (2, 1, 2, 3, 0, 0.8), (3, 0, 8, 1, -1, 0.4), (3, 8, 16, 2, 1, 0.6),
    # high pileup: gathers of several chunks and the incremental tile-minimum refresh (event 29,
    # N = 5194: about 150 members per gather); for kt (event 17, N = 9867, gathers beyond the
    # four chunks a recomputation keeps in registers) a jet that takes the merged jet as its
    # neighbour can get a larger diJ
    (4, 29, 30, 0, -1, 0.4), (4, 17, 18, 0, 1, 0.4),
])
def test_generated_events(emu_bin, workload, e0, e1, variant, p, R):
    out = run(emu_bin, workload, e0, e1, variant, p, R, 0)
    assert "done: 0 mismatches" in out


def _edge_events():
    base = np.array([[3.0, 1.0, 0.5, 4.0], [0.5, 2.0, -1.0, 3.0], [-1.0, -1.0, 0.2, 2.0]])
    lat = []
    for i in range(12):
        for j in range(12):
            eta, phi = -1.0 + 0.25 * i, 0.25 * j
            pt = 1.0 + 0.125 * ((i * 7 + j * 3) % 5)
            lat.append([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)])
    return {
        "duplicates": np.repeat(base, 40, axis=0),
        "zero_pt": np.array([[0.0, 0.0, 5.0, 5.0], [0.0, 0.0, -7.0, 7.0], [1.0, 1.0, 0.0, 2.0],
                             [0.0, 0.0, 1.0, 2.0], [1e-3, 0.0, 0.0, 1e-3], [0.0, 0.0, 0.0, 0.0]]),
        "lattice": np.array(lat),
        "single": np.array([[1.0, 2.0, 3.0, 10.0]]),
    }


@pytest.mark.parametrize("name", ["duplicates", "zero_pt", "lattice", "single"])
@pytest.mark.parametrize("variant", [0, 3])
@pytest.mark.parametrize("p,R", [(1, 0.7), (0, 0.7), (-1, 0.3)])
def test_edge_events(emu_bin, tmp_path, name, variant, p, R):
    path = tmp_path / f"{name}.bin"
    _edge_events()[name].astype("<f8").tofile(path)
    out = run(emu_bin, "file", path, variant, p, R)
    assert " ok" in out and "MISMATCH" not in out


@pytest.mark.parametrize("seed", [1, 2])
def test_kernel_kinematics_bit_identical(emu_bin, seed):
    """The kernel's straight-line common kinematics path and its fall-back give
    the same bits as the shared scalar fjmi_rap/fjmi_phi/fjmi_weight, on edge
    cases (zeros, subnormals, infinities, NaN, E == |pz|, underflowing kt^2)
    and on random momenta over 160 binary decades."""
    out = run(emu_bin, "kin", 100000, seed)
    assert "kinematics: 0 of" in out


@pytest.mark.parametrize("workload,e0,e1,variant,p,R", [(3, 0, 6, 0, -1, 0.4), (3, 0, 6, 1, 1, 0.4)])
def test_pool_exhaustion_rebuilds(emu_bin, workload, e0, e1, variant, p, R):
    """A tile pool of exactly N entries (no room for relocated segments): every
    exhaustion lays all segments out again, and the results stay bit-exact."""
    out = run(emu_bin, workload, e0, e1, variant, p, R, 0, 1 << 30, 0, 1, env=TIGHT_ENV)
    assert "done: 0 mismatches" in out
    rebuilds = int(out.split("rebuilds:")[1].split()[0])
    assert rebuilds > 0, out[-500:]


def test_queue_index_bounds(emu_bin):
    """A queue head yielding an index below 0 or at/after the queue's end (the
    one-event form pops entry 0 only) leaves the launch without writing."""
    out = run(emu_bin, "qbound")
    assert "done: 0 mismatches" in out and out.count(": ok") == 3, out


@pytest.mark.parametrize("variant,p", [(1, -1), (0, 1)])
def test_headless_single_pass(emu_bin, variant, p):
    """Launches without a queue head (how the one-event entry points launch):
    one pass over queue entry 0, bit-exact."""
    out = run(emu_bin, 3, 20, 24, variant, p, 0.4, env={"FJMI_EMU_HEADLESS": "1"})
    assert "done: 0 mismatches" in out


@pytest.mark.parametrize("workload,e0,e1,variant,p,R", [(3, 0, 4, 1, -1, 0.4), (4, 29, 30, 0, 1, 0.4)])
def test_gather_counter_restarts(emu_bin, workload, e0, e1, variant, p, R):
    """Restarting the per-event gather counter (all stamps cleared) keeps every
    result bit-exact, single-chunk and several-chunk gathers alike."""
    out = run(emu_bin, workload, e0, e1, variant, p, R, 0, env=STAMP_ENV)
    assert "done: 0 mismatches" in out


@pytest.mark.parametrize("workload,e0,e1,variant,p,scheme,R", [
    (1, 0, 1, 0, -1.0, 0, 0.4), (1, 0, 1, 3, 1.0, 0, 0.4),          # standard definitions on the general kernel
    (1, 0, 1, 1, 0.5, 1, 0.7), (2, 0, 1, 0, 2.0, 2, 1.0), (3, 0, 3, 1, -1.0, 3, 0.4),
    (3, 3, 6, 2, 0.3, 4, 0.6), (1, 0, 1, 3, -0.5, 5, 0.4), (1, 0, 1, 1, 1.0, 6, 0.4),
    (3, 6, 9, 3, -2.5, 7, 0.5), (3, 9, 12, 0, 0.0, 8, 0.8),
    (4, 29, 30, 0, 0.5, 1, 0.4),                                    # several-chunk gathers (N = 5194)
])
def test_general_definitions(emu_bin, workload, e0, e1, variant, p, scheme, R):
    """tiled_gen_kernel (any real genkt exponent, FastJet's recombination schemes)
    bit-exact against the oracle's restatement of the same definition."""
    out = run(emu_bin, "gen", workload, e0, e1, variant, p, scheme, R)
    assert "done: 0 mismatches" in out and "MISMATCH" not in out


@pytest.mark.parametrize("workload,e0,e1,variant,p,R", [
    (3, 0, 4, 0, -1, 1.0), (3, 4, 8, 1, 1, 1.0), (2, 0, 1, 1, 0, 1.5), (1, 0, 1, 0, -1, 2.0),
])
def test_cooperative_kernel(emu_bin, workload, e0, e1, variant, p, R):
    """tiled_coop_kernel (8 wavefronts per event sharing gathers of several
    chunks, emulated with workgroup barriers): bit-exact against the oracle."""
    out = run(emu_bin, workload, e0, e1, variant, p, R, 0, env={"FJMI_EMU_COOP": "1"})
    assert "done: 0 mismatches" in out
