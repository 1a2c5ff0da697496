"""Restatement of the reference's file formats in numpy -- TEST INFRASTRUCTURE.

Only tests/ may import this module, as the checker of the library's readers
and writers (thx_thu_*, thx_mrc_*, thx_ctx_load_thu).

* MRC (ImageFile, src/Image/ImageFile.cpp:174-300, MRCHeader in
  include/Image/ImageFile.h): a 1024-byte header of 256 little-endian 32-bit
  words (nx, ny, nz, mode at words 0-3; mx, my, mz 7-9; cell lengths 10-12 and
  angles 13-15 as floats; mapc, mapr, maps 16-18; ispg 22; nsymbt 23; "MAP "
  at byte 208), nsymbt bytes of extended header, then the slices, x fastest,
  in mode 0 (int8), 1 (int16) or 2 (float32).
* .thu (Database, src/Database.cpp:100-300; fields include/Database.h:12-88):
  one particle per line, 27 whitespace-separated fields: voltage, defocusU,
  defocusV, defocusTheta, Cs, amplitudeContrast, phaseShift, particle path,
  micrograph path, coordinate x, y, group id, class id, quaternion (4),
  k1 k2 k3, translation x y, translation deviations x y, defocus factor, its
  deviation, score.  STAR_2_THU.py (script/STAR_2_THU.py:62-118) writes the
  input form (%18.6f), Optimiser::saveDatabase (src/Optimiser.cpp:8167-8268)
  the output form (%18.9lf).
"""
import numpy as np

MODE_DTYPE = {0: np.int8, 1: np.int16, 2: np.float32}


def mrc_write(path, data, mode=2, ext=b"", cell=None):
    """An MRC file of data [nz][ny][nx] in the given mode with an optional
    extended header."""
    data = np.asarray(data)
    if data.ndim == 2:
        data = data[None]
    nz, ny, nx = data.shape
    w = np.zeros(256, np.int32)
    w[0:4] = (nx, ny, nz, mode)
    w[7:10] = (nx, ny, nz)
    f = w.view(np.float32)
    f[10:13] = cell if cell is not None else (nx, ny, nz)
    f[13:16] = 90.0
    w[16:19] = (1, 2, 3)
    w[22] = 1
    w[23] = len(ext)
    head = bytearray(w.tobytes())
    head[208:212] = b"MAP "
    with open(path, "wb") as fh:
        fh.write(bytes(head))
        fh.write(ext)
        fh.write(np.ascontiguousarray(data, MODE_DTYPE[mode]).tobytes())


def mrc_read(path):
    """(header words as int32, header as float32, data [nz][ny][nx] float64)."""
    raw = open(path, "rb").read()
    w = np.frombuffer(raw[:1024], np.int32)
    f = np.frombuffer(raw[:1024], np.float32)
    nx, ny, nz, mode = (int(x) for x in w[:4])
    off = 1024 + int(w[23])
    dt = MODE_DTYPE[mode]
    n = nx * ny * nz
    data = np.frombuffer(raw[off:off + n * np.dtype(dt).itemsize], dt).reshape(nz, ny, nx)
    return w, f, raw[208:212], data.astype(np.float64)


THU_FIELDS = 27


def thu_line_input(ctf, particle, micrograph, coord, group):
    """A line as STAR_2_THU.py writes it (%18.6f; class, pose and deviations
    zero, defocus factor 1)."""
    v = list(ctf) + [particle, micrograph, coord[0], coord[1], group, 0] + [0.0] * 4 + \
        [0.0] * 3 + [0.0] * 4 + [1.0, 0.0, 0.0]
    out = []
    for i, x in enumerate(v):
        if i in (7, 8):
            out.append(str(x))
        elif i in (11, 12):
            out.append(str(int(x)))
        else:
            out.append(f"{x:18.6f}")
    return "    ".join(out) + "\n"


def thu_parse(path):
    """The lines of a .thu file split into their 27 fields (blank lines skipped)."""
    rows = []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        assert len(t) == THU_FIELDS, len(t)
        rows.append(t)
    return rows
