"""The drop-in command-line binaries (fastjet_gpu_amd/bin/{tri_matrix,n_array,grid,
tri_matrix_empty}): argument handling on the CPU, and on the GPU the printed jets
against the oracle's, formatted like the reference driver (main.cc:87-116, 339-367).
"""
import ctypes
import ctypes.util
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "fastjet_gpu_amd", "bin")
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
for _f in ("hypot", "atan2"):
    getattr(_libm, _f).argtypes = [ctypes.c_double, ctypes.c_double]
    getattr(_libm, _f).restype = ctypes.c_double
_libm.log.argtypes = [ctypes.c_double]
_libm.log.restype = ctypes.c_double


def run(binary, args, stdin=None, timeout=120):
    path = os.path.join(BIN, binary)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    return subprocess.run([path, *args], input=stdin, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "Unrecognized option --bogus"),
    (["-r"], "Missing argument to option -r"),
    (["-r", "abc"], "Error while parsing argument to option -r"),
    (["--ptmin", "-1"], "Error while parsing argument to option --ptmin"),
    (["--repeat", "x"], "Error while parsing argument to option --repeat"),
    (["-combine"], "Missing argument to option -combine"),
    (["-f"], "Missing argument to option -f"),
])
def test_argument_errors(args, msg):
    """Bad flags exit with status 1 and the reference's message, before any GPU work."""
    r = run("tri_matrix", args)
    assert r.returncode == 1
    assert r.stderr.strip() == msg


def expected_lines(events, p, R, ptmin, cartesian=False):
    """The jet listing main.cc prints for one clustering of the concatenated events."""
    ev = np.concatenate(events)
    h1, h2, hd, jets = orc.n2plain(ev, p, R)
    keep = [j for j in jets if j[0] * j[0] + j[1] * j[1] >= ptmin * ptmin]
    keep.sort(key=lambda j: -(j[0] * j[0] + j[1] * j[1]))
    out = []
    for k, (px, py, pz, E) in enumerate(keep):
        if cartesian:
            out.append(f"{k:5d}{px:16.8f}{py:16.8f}{pz:16.8f}{E:16.8f}")
        else:
            pt = _libm.hypot(px, py)
            phi = _libm.atan2(py, px)
            while phi > 2 * np.pi:
                phi -= 2 * np.pi
            while phi < 0.0:
                phi += 2 * np.pi
            m2 = max(0.0, (E + pz) * (E - pz) - pt * pt)
            epz = E + abs(pz)
            y = 0.5 * _libm.log((pt * pt + m2) / (epz * epz))
            if pz > 0:
                y = -y
            out.append(f"{k:5d}{y:16.8f}{phi:16.8f}{pt:16.8f}")
    return len(ev), out


def jet_lines(stdout):
    return [l for l in stdout.splitlines() if re.match(r"^\s{0,4}\d+\s+-?\d+\.\d{8}", l)]


@pytest.mark.gpu
@pytest.mark.parametrize("binary", ["tri_matrix", "n_array", "grid"])
@pytest.mark.parametrize("flag,p", [("-antikt", -1), ("-cam", 0), ("-kt", 1)])
def test_reference_run_targets(binary, flag, p):
    """The reference's own validation runs (Makefile:69-87): R 0.4, ptmin 1.0 on 808_cms.dat."""
    path = os.path.join(GOLDEN, "808_cms.dat")
    print(f"[gpu cli] {binary} {flag} -r 0.4 -p 1.0 < 808_cms.dat", flush=True)
    with open(path) as fh:
        r = run(binary, [flag, "-r", "0.4", "-p", "1.0"], stdin=fh.read())
    assert r.returncode == 0, r.stderr
    n, want = expected_lines(orc.read_events(path)[:1], p, 0.4, 1.0)
    assert jet_lines(r.stdout) == want
    assert f"clustered {n} particles into {len(want)} jets above 1 GeV in " in r.stdout
    assert r.stdout.splitlines()[0].startswith("Running on HIP device")


@pytest.mark.gpu
def test_csv_cartesian_combine_repeat():
    path = os.path.join(GOLDEN, "multi-event.dat")
    events = orc.read_events(path)
    print("[gpu cli] grid --csv/--cartesian/--combine", flush=True)
    r = run("grid", ["-antikt", "-r", "0.7", "--cartesian", "--combine", "2", "--repeat", "3", "-f", path])
    assert r.returncode == 0, r.stderr
    got = jet_lines(r.stdout)
    want = []
    for k in range(0, len(events), 2):
        want += expected_lines(events[k:k + 2], -1, 0.7, 0.0, cartesian=True)[1]
    assert got == want
    assert r.stdout.count("found ") == 5 + 5 * 3  # "found N particles in 2 events" + one line per repetition
    assert "in 2 events" in r.stdout
    assert re.search(r"clustered \d+ particles into \d+ jets above 0 GeV in \d+\.\d+ \+/- \d+\.\d+ ms", r.stdout)
    rc = run("tri_matrix", ["-cam", "-r", "0.4", "--csv", "-f", path])
    assert rc.returncode == 0, rc.stderr
    csv = [l for l in rc.stdout.splitlines() if re.match(r"^\d+,\d+,\d+\.\d$", l)]
    assert len(csv) == len(events)
    for line, ev in zip(csv, events):
        n, want = expected_lines([ev], 0, 0.4, 0.0)
        assert line.startswith(f"{n},{len(want)},")


@pytest.mark.gpu
def test_empty_engine_runs():
    """tri_matrix_empty: the per-step launch schedule with empty kernels (no result)."""
    path = os.path.join(GOLDEN, "354_fastjet.dat")
    print("[gpu cli] tri_matrix_empty", flush=True)
    r = run("tri_matrix_empty", ["-antikt", "-r", "0.4", "--csv", "-f", path])
    assert r.returncode == 0, r.stderr
    assert re.search(r"^354,\d+,\d+\.\d$", r.stdout, re.M)


def _g(x, prec):
    """std::cout << x under std::defaultfloat with precision prec (C's %g)."""
    return f"{x:.{max(prec, 1)}g}"


def expected_stdout(groups, p, R, ptmin, cartesian=False, csv=False, repeat=1, combine=1):
    """The whole stdout of main.cc:269-368 after the banner, with the timings as
    <t> and <s>.  Keeps the reference's stream state: print_jets leaves precision
    8 set (main.cc:87), which the next groups' "found ... ms" lines inherit."""
    out, prec = [], 6
    for grp in groups:
        n, jl = expected_lines(grp, p, R, ptmin, cartesian)
        if not csv:
            k = len(grp)
            out.append(f"found {n} particles" + (f" in {k} event{'' if k == 1 else 's'}" if combine != 1 else ""))
            for _ in range(repeat):
                out.append(f"found {len(jl)} jets above {_g(ptmin, prec)} GeV in <t> ms" if ptmin > 0
                           else f"found {len(jl)} jets in <t> ms")
            out += jl + [""]
            prec = 8
            out.append(f"clustered {n} particles into {len(jl)} jets above {_g(ptmin, prec)} GeV in <t>"
                       + (" +/- <s>" if repeat > 1 else "") + " ms")
        else:
            out.append(f"{n},{len(jl)},<t>" + (",<s>" if repeat > 1 else ""))
    return out


def normalised(stdout, csv=False, repeat=1):
    """stdout without the 3 banner lines, timings replaced by <t> / <s>; the mean and
    sigma of several repetitions must share the fixed precision main.cc:350-358 picks."""
    lines = stdout.splitlines()
    assert lines[0].startswith("Running on HIP device ")
    assert re.fullmatch(r"  - maximum shared memory per block: \d+ kB", lines[1])
    assert re.fullmatch(r"  - kernel printf buffer size:       \d+ kB", lines[2])
    out = []
    for l in lines[3:]:
        m = re.fullmatch(r"(.* in )(\d+(?:\.\d+)?) \+/- (\d+(?:\.\d+)?) ms", l)
        c = re.fullmatch(r"(\d+),(\d+),(\d+(?:\.\d+)?)(?:,(\d+(?:\.\d+)?))?", l) if csv else None
        if m:
            a, b = m.group(2), m.group(3)
            assert len(a.partition(".")[2]) == len(b.partition(".")[2]), l
            out.append(m.group(1) + "<t> +/- <s> ms")
        elif c:
            if c.group(4) is not None:
                assert len(c.group(3).partition(".")[2]) == len(c.group(4).partition(".")[2]), l
            elif repeat == 1:
                assert re.fullmatch(r"\d+\.\d", c.group(3)), l  # precision 1 for a single repetition
            out.append(f"{c.group(1)},{c.group(2)},<t>" + (",<s>" if c.group(4) is not None else ""))
        else:
            out.append(re.sub(r" in \S+ ms$", " in <t> ms", l))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("binary,args,fname,groups,p,R,ptmin,cart,csv,rep,comb", [
    ("tri_matrix", ["-antikt", "-r", "0.4", "-p", "1.0"], "808_cms.dat", None, -1, 0.4, 1.0, False, False, 1, 1),
    ("grid", ["-antikt", "-r", "0.7", "--cartesian", "--combine", "2", "--repeat", "3"], "multi-event.dat", 2,
     -1, 0.7, 0.0, True, False, 3, 2),
    ("n_array", ["-kt", "-r", "1.0", "-p", "0.5", "--repeat", "2"], "354_fastjet.dat", None, 1, 1.0, 0.5, False,
     False, 2, 1),
    ("tri_matrix", ["-cam", "-r", "0.4", "--csv"], "multi-event.dat", 1, 0, 0.4, 0.0, False, True, 1, 1),
    ("grid", ["-kt", "-r", "0.4", "-p", "2.25", "--combine", "3"], "multi-event.dat", 3, 1, 0.4, 2.25, False,
     False, 1, 3),
])
def test_whole_stdout(binary, args, fname, groups, p, R, ptmin, cart, csv, rep, comb):
    """Every stdout line of the CLI equals main.cc's output for the same input, the
    timings aside: banner, per-event and per-repetition lines, jets, blank lines,
    summary or CSV, stream precision carried across event groups."""
    path = os.path.join(GOLDEN, fname)
    events = orc.read_events(path)
    grp = [events[k:k + (groups or 1)] for k in range(0, len(events), groups or 1)]
    print(f"[gpu cli] {binary} {' '.join(args)} -f {fname}", flush=True)
    r = run(binary, [*args, "-f", path])
    assert r.returncode == 0, r.stderr
    want = expected_stdout(grp, p, R, ptmin, cart, csv, rep, comb)
    got = normalised(r.stdout, csv, rep)
    assert got == want, next(((k, g, w) for k, (g, w) in enumerate(zip(got, want)) if g != w), (len(got), len(want)))
