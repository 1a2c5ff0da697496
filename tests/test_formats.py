"""The .thu and MRC readers and writers of the library (host code, no GPU)
against the numpy restatement of the formats (oracle/formats.py): files
written by one side are read by the other, field by field and value by
value, in every MRC mode the reference reads (int8, int16, float32) and with an
extended header; errors name the line or the file."""
import ctypes as C
import os

import numpy as np
import pytest

import formats
import thx


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mrc_read_matches_format(tmp_path, mode):
    rng = np.random.default_rng(mode)
    data = rng.normal(0, 20, size=(5, 10, 12))
    if mode != 2:
        data = np.round(data).clip(-100, 100)
    p = str(tmp_path / f"s{mode}.mrcs")
    formats.mrc_write(p, data, mode=mode, ext=bytes(range(160)))
    assert thx.mrc_info(p) == (12, 10, 5, mode)
    got = thx.mrc_read(p)
    np.testing.assert_array_equal(got, data.astype(formats.MODE_DTYPE[mode]).astype(np.float32))
    np.testing.assert_array_equal(thx.mrc_read(p, 2, 2), got[2:4])


def test_mrc_write_matches_format(tmp_path):
    rng = np.random.default_rng(9)
    data = rng.normal(size=(3, 16, 16)).astype(np.float32)
    p = str(tmp_path / "w.mrc")
    thx.mrc_write(p, data, pixel_size=1.5)
    w, f, stamp, back = formats.mrc_read(p)
    assert tuple(w[:4]) == (16, 16, 3, 2) and tuple(w[7:10]) == (16, 16, 3)
    assert tuple(w[16:19]) == (1, 2, 3) and w[22] == 1 and w[23] == 0
    np.testing.assert_allclose(f[10:13], [24.0, 24.0, 4.5])
    np.testing.assert_array_equal(f[13:16], 90.0)
    assert stamp == b"MAP "
    assert f[19] == data.min() and f[20] == data.max()
    np.testing.assert_array_equal(back, data)


def test_mrc_errors(tmp_path):
    p = str(tmp_path / "bad.mrc")
    formats.mrc_write(p, np.zeros((2, 4, 4)), mode=2)
    with pytest.raises(thx.THXError, match="out of range"):
        thx.mrc_read(p, 1, 2)
    raw = bytearray(open(p, "rb").read())
    raw[12:16] = np.int32(6).tobytes()          # mode 6 (uint16) is not read
    open(p, "wb").write(bytes(raw))
    with pytest.raises(thx.THXError, match="mode"):
        thx.mrc_info(p)
    with pytest.raises(thx.THXError, match="cannot open"):
        thx.mrc_info(str(tmp_path / "missing.mrc"))


def _input_thu(path, n, rng):
    rows = []
    with open(path, "w") as fh:
        for i in range(n):
            ctf = [300e3, rng.uniform(1e4, 3e4), rng.uniform(1e4, 3e4), rng.uniform(0, 3),
                   2.7e7, 0.07, 0.0]
            rows.append((ctf, f"{i + 1:06d}@stack_{i % 3}.mrcs", f"mic_{i // 4}.mrc",
                         (rng.uniform(0, 4096), rng.uniform(0, 4096)), 1 + i % 2))
            fh.write(formats.thu_line_input(*rows[-1]))
            if i == 3:
                fh.write("\n   \t\n")             # blank lines are skipped
    return rows


def test_thu_read_matches_format(tmp_path):
    rng = np.random.default_rng(3)
    p = str(tmp_path / "in.thu")
    rows = _input_thu(p, 9, rng)
    recs = thx.thu_read(p)
    assert len(recs) == 9
    for r, (ctf, part, mic, coord, grp) in zip(recs, rows):
        np.testing.assert_array_equal(np.array(r.ctf),
                                      np.array([float(f"{x:18.6f}") for x in ctf], np.float32))
        assert r.particle.decode() == part and r.micrograph.decode() == mic
        np.testing.assert_allclose(list(r.coord), [float(f"{c:18.6f}") for c in coord])
        assert r.group == grp and r.cls == 0 and r.df == 1.0 and list(r.quat) == [0.0] * 4


def test_thu_write_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    src = str(tmp_path / "in.thu")
    _input_thu(src, 5, rng)
    recs = thx.thu_read(src)
    for i, r in enumerate(recs):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        r.quat[:] = q
        r.k[:] = rng.uniform(0.001, 0.1, 3)
        r.trans[:] = rng.uniform(-3, 3, 2)
        r.stdTrans[:] = rng.uniform(0.1, 1, 2)
        r.cls = i % 2
        r.score = rng.uniform(0, 1)
    out = str(tmp_path / "out.thu")
    thx.thu_write(out, recs[:3])
    thx.thu_write(out, recs[3:], append=True)
    rows = formats.thu_parse(out)
    assert len(rows) == 5
    for r, t in zip(recs, rows):
        assert t[7] == r.particle.decode() and t[8] == r.micrograph.decode()
        assert int(t[11]) == r.group and int(t[12]) == r.cls
        np.testing.assert_allclose([float(x) for x in t[13:17]], list(r.quat), atol=1e-9)
        np.testing.assert_allclose([float(x) for x in t[17:20]], list(r.k), atol=1e-9)
        np.testing.assert_allclose([float(x) for x in t[20:24]],
                                   list(r.trans) + list(r.stdTrans), atol=1e-9)
        assert float(t[26]) == pytest.approx(r.score, abs=1e-9)
    back = thx.thu_read(out)
    for a, b in zip(recs, back):
        np.testing.assert_allclose(list(b.quat), list(a.quat), atol=1e-9)
        np.testing.assert_array_equal(list(b.ctf), list(a.ctf))


def test_thu_short_line_is_an_error(tmp_path):
    p = str(tmp_path / "bad.thu")
    rng = np.random.default_rng(5)
    _input_thu(p, 2, rng)
    with open(p, "a") as fh:
        fh.write("300000.0 1.0 2.0\n")
    with pytest.raises(thx.THXError, match="line 3 .* has 3 fields"):
        thx.thu_read(p)
