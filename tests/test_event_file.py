"""Binary event-batch files (fjmi_events_write / _info / _read in include/fjmi.h) and the text reader for the
reference driver's event format (main.cc:34-70).  Host code only: no GPU is touched."""
import os
import struct

import numpy as np
import pytest

import fastjet_gpu_amd as fj
from tests.conftest import GOLDEN
from oracle import oracle as orc


def _batch(seed=3, ne=7, with_empty=True):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 40, ne)
    if with_empty:
        sizes[2] = 0
    off = np.zeros(ne + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    p4 = rng.normal(size=(int(off[-1]), 4))
    return p4, off


@pytest.mark.parametrize("with_p,with_R", [(False, False), (True, False), (False, True), (True, True)])
def test_round_trip(tmp_path, with_p, with_R):
    p4, off = _batch()
    ne = len(off) - 1
    p = np.array([(-1, 0, 1)[e % 3] for e in range(ne)], np.int32) if with_p else None
    R = np.linspace(0.2, 1.5, ne) if with_R else None
    path = str(tmp_path / "ev.fjmi")
    fj.write_events(path, p4, off, p, R)
    assert os.path.getsize(path) == 32 + 8 * (ne + 1) + 32 * len(p4) + (4 * ne if with_p else 0) + (8 * ne if with_R else 0)
    q4, qoff, qp, qR = fj.read_events(path)
    assert q4.tobytes() == p4.tobytes()
    assert qoff.tobytes() == off.tobytes()
    assert (qp is None) == (not with_p) and (qR is None) == (not with_R)
    if with_p:
        assert qp.tobytes() == p.tobytes()
    if with_R:
        assert qR.tobytes() == R.tobytes()
    assert not os.path.exists(path + ".tmp")


def test_layout_is_the_documented_one(tmp_path):
    p4 = np.arange(8, dtype=np.float64).reshape(2, 4)
    off = np.array([0, 2], np.int64)
    path = str(tmp_path / "one.fjmi")
    fj.write_events(path, p4, off, np.array([1], np.int32), np.array([0.7]))
    raw = open(path, "rb").read()
    assert raw[:8] == b"FJMIEVT1"
    assert struct.unpack("<QQII", raw[8:32]) == (1, 2, 3, 0)
    assert struct.unpack("<2q", raw[32:48]) == (0, 2)
    assert struct.unpack("<8d", raw[48:112]) == tuple(range(8))
    assert struct.unpack("<i", raw[112:116]) == (1,)
    assert struct.unpack("<d", raw[116:124]) == (0.7,)
    assert len(raw) == 124


def test_zero_events(tmp_path):
    path = str(tmp_path / "empty.fjmi")
    fj.write_events(path, np.zeros((0, 4)), np.zeros(1, np.int64))
    q4, qoff, qp, qR = fj.read_events(path)
    assert len(q4) == 0 and qoff.tolist() == [0] and qp is None and qR is None


def _corrupt(tmp_path, name, mutate):
    p4, off = _batch()
    ne = len(off) - 1
    path = str(tmp_path / name)
    fj.write_events(path, p4, off, np.zeros(ne, np.int32), np.full(ne, 0.4))
    raw = bytearray(open(path, "rb").read())
    raw = mutate(raw, ne, len(p4))
    with open(path, "wb") as fh:
        fh.write(bytes(raw))
    return path


def _p_at(ne, np_):
    return 32 + 8 * (ne + 1) + 32 * np_


@pytest.mark.parametrize("case,mutate,msg", [
    ("magic", lambda r, ne, np_: b"FJMIEVT2" + r[8:], "bad magic"),
    ("truncated", lambda r, ne, np_: r[:-1], "bytes where the header implies"),
    ("trailing", lambda r, ne, np_: r + b"\0", "bytes where the header implies"),
    ("header", lambda r, ne, np_: r[:20], "truncated header"),
    ("flags", lambda r, ne, np_: r[:24] + struct.pack("<I", 7) + r[28:], "unknown flags"),
    ("offset0", lambda r, ne, np_: r[:32] + struct.pack("<q", 1) + r[40:], "bad offsets"),
    ("monotonic", lambda r, ne, np_: r[:48] + struct.pack("<q", 10 ** 6) + r[56:], "not monotonic"),
    ("exponent", lambda r, ne, np_: r[:_p_at(ne, np_)] + struct.pack("<i", 2) + r[_p_at(ne, np_) + 4:], "exponent 2"),
    ("radius", lambda r, ne, np_: r[:-8] + struct.pack("<d", 0.0), "radius of event"),
    ("radius_nan", lambda r, ne, np_: r[:-8] + struct.pack("<d", float("nan")), "radius of event"),
])
def test_corrupt_files_are_rejected(tmp_path, case, mutate, msg):
    path = _corrupt(tmp_path, case + ".fjmi", mutate)
    with pytest.raises(fj.FjmiError, match=msg):
        fj.read_events(path)


def test_missing_file(tmp_path):
    with pytest.raises(fj.FjmiError, match="cannot open"):
        fj.read_events(str(tmp_path / "nope.fjmi"))


def test_writer_rejects_bad_offsets(tmp_path):
    path = str(tmp_path / "bad.fjmi")
    with pytest.raises(fj.FjmiError, match="monotonic"):
        fj.write_events(path, np.zeros((3, 4)), np.array([0, 2, 1, 3], np.int64))
    with pytest.raises(fj.FjmiError, match="offsets"):
        fj.write_events(path, np.zeros((3, 4)), np.array([1, 3], np.int64))
    assert not os.path.exists(path) and not os.path.exists(path + ".tmp")


def test_text_events_match_the_reference_format():
    """The golden text files read to the same events the oracle's reader gives; each text file converts to a
    binary file losslessly (the values are parsed once, then copied bit for bit)."""
    for name in ("354_fastjet.dat", "808_cms.dat", "multi-event.dat"):
        p4, off = fj.read_text_events(os.path.join(GOLDEN, name))
        ref = orc.read_events(os.path.join(GOLDEN, name))
        assert len(off) - 1 == len(ref)
        for e, ev in enumerate(ref):
            assert p4[off[e]:off[e + 1]].tobytes() == ev.tobytes()
    assert np.diff(fj.read_text_events(os.path.join(GOLDEN, "808_cms.dat"))[1]).tolist() == [808]


def test_text_reader_errors_like_the_reference(tmp_path):
    good = tmp_path / "good.dat"
    good.write_text("# header\n1 2 3 4\n5 6 7 8 extra\n\n#\n9 10 11 12\n")
    p4, off = fj.read_text_events(str(good))
    assert off.tolist() == [0, 2, 3] and p4[1].tolist() == [5, 6, 7, 8]
    bad = tmp_path / "bad.dat"
    bad.write_text("1 2 3 4\n1 2 3\n")
    with pytest.raises(ValueError, match="Error while parsing particles"):
        fj.read_text_events(str(bad))


def test_text_to_binary_to_oracle(tmp_path):
    """A text file converted to binary clusters to the same history as the text events themselves."""
    p4, off = fj.read_text_events(os.path.join(GOLDEN, "multi-event.dat"))
    path = str(tmp_path / "multi.fjmi")
    fj.write_events(path, p4, off)
    q4, qoff, _, _ = fj.read_events(path)
    a = orc.cluster_batch(p4, off, -1, 0.4, 2)
    b = orc.cluster_batch(q4, qoff, -1, 0.4, 2)
    for x, y in zip(a[:5], b[:5]):
        assert x.tobytes() == y.tobytes()
