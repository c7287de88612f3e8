"""Reader and writer of the reference's restart files (restflds / restprtl).

Each file is one Fortran-unformatted sequential record (code/output.F90:
2200-2223, read back by code/restart.F90:220-247):

  restflds   mx, my, mz (int32); ex, ey, ez, bx, by, bz (float32, mx*my*mz
             each, x fastest, ghost cells included); dseed (float64);
             lap (int32); xinject, xinject2, xinject3 (float64);
             leftwall (float32); walloc (float64)
  restprtl   ions, lecs, maxptl, maxhlf, totalpartnum (int32); then x, y, z,
             u, v, w, ch (float32) and ind, proc, splitlev (int32), each as
             the ions followed by the electrons

A record is framed by its byte length as a 4-byte integer before and after
it.  A record longer than 2**31 - 9 bytes is split into subrecords of at most
that many bytes (the convention of gfortran and Intel Fortran): the leading
length of a subrecord is negative when another subrecord follows, the
trailing length negative when another precedes.  The reader also takes the
single record whose two lengths hold its byte count modulo 2**32, the form
LLVM flang's runtime writes beyond 2 GiB (readable by that runtime only).

Used for fixture exchange with the reference and to inspect the Fortran
host's checkpoints; the library itself does not read these files.
"""
import os

import numpy as np

_MAXSUB = 2 ** 31 - 9

FIELD_NAMES = ("ex", "ey", "ez", "bx", "by", "bz")
PART_REALS = ("x", "y", "z", "u", "v", "w", "ch")
PART_INTS = ("ind", "proc", "splitlev")


def _pieces(mm, path):
    """(offset, length) of the payload pieces of the file's one record"""
    size = len(mm)
    pieces, pos = [], 0
    while pos + 8 <= size:
        n = int(np.frombuffer(mm, "<i4", 1, pos)[0])
        length = abs(n)
        end = pos + 4 + length
        if end + 4 > size or abs(int(np.frombuffer(mm, "<i4", 1, end)[0])) != length:
            break
        pieces.append((pos + 4, length))
        pos = end + 4
        if n >= 0:
            return pieces
    # one record whose markers hold its length modulo 2**32
    if size >= 8:
        lead = int(np.frombuffer(mm, "<i4", 1, 0)[0])
        tail = int(np.frombuffer(mm, "<i4", 1, size - 4)[0])
        if lead == tail and (size - 8) % 2 ** 32 == lead % 2 ** 32:
            return [(4, size - 8)]
    raise ValueError(f"{path}: not a sequential unformatted record")


def read_record(path):
    """the first record of a sequential unformatted file, as a bytes-like
    object (a view of the memory-mapped file when the record is one piece)"""
    import mmap
    with open(path, "rb") as f:
        if os.fstat(f.fileno()).st_size == 0:
            raise ValueError(f"{path}: empty file")
        mm = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
    pieces = _pieces(mm, path)
    if len(pieces) == 1:
        off, length = pieces[0]
        return memoryview(mm)[off:off + length]
    return b"".join(mm[off:off + length] for off, length in pieces)


def write_record(path, payload):
    """one sequential unformatted record holding payload (bytes-like, or a
    list of bytes-like parts written in order)"""
    parts = payload if isinstance(payload, (list, tuple)) else [payload]
    views = [memoryview(p).cast("B") for p in parts]
    n = sum(len(v) for v in views)
    with open(path, "wb") as f:
        if n == 0:
            f.write(np.int32(0).tobytes() * 2)
            return
        done, first = 0, True
        it = iter(views)
        cur, cpos = next(it), 0
        while done < n:
            size = min(_MAXSUB, n - done)
            f.write(np.int32(-size if done + size < n else size).tobytes())
            left = size
            while left:
                while cpos == len(cur):
                    cur, cpos = next(it), 0
                take = min(left, len(cur) - cpos)
                f.write(cur[cpos:cpos + take])
                cpos += take
                left -= take
            done += size
            f.write(np.int32(size if first else -size).tobytes())
            first = False


class _Cursor:
    def __init__(self, buf, path):
        self.buf, self.pos, self.path = buf, 0, path

    def take(self, dtype, count=1):
        dt = np.dtype(dtype)
        nbytes = dt.itemsize * count
        if self.pos + nbytes > len(self.buf):
            raise ValueError(f"{self.path}: record too short ({len(self.buf)} bytes) for its contents")
        a = np.frombuffer(self.buf, dt, count, self.pos).copy()
        self.pos += nbytes
        return a

    def done(self):
        if self.pos != len(self.buf):
            raise ValueError(f"{self.path}: {len(self.buf) - self.pos} bytes left over in the record")


def read_fields(path):
    """dict with mx, my, mz, the six fields as (mz, my, mx) float32 arrays,
    dseed, lap, xinject, xinject2, xinject3, leftwall, walloc"""
    c = _Cursor(read_record(path), path)
    mx, my, mz = (int(v) for v in c.take("<i4", 3))
    if min(mx, my, mz) <= 0:
        raise ValueError(f"{path}: grid extent {mx} x {my} x {mz}")
    n = mx * my * mz
    out = dict(mx=mx, my=my, mz=mz)
    for k in FIELD_NAMES:
        out[k] = c.take("<f4", n).reshape(mz, my, mx)
    out["dseed"] = float(c.take("<f8")[0])
    out["lap"] = int(c.take("<i4")[0])
    for k in ("xinject", "xinject2", "xinject3"):
        out[k] = float(c.take("<f8")[0])
    out["leftwall"] = float(c.take("<f4")[0])
    out["walloc"] = float(c.take("<f8")[0])
    c.done()
    return out


def write_fields(path, fields, lap, dseed=123457.0, xinject=0.0, xinject2=0.0, xinject3=0.0, leftwall=15.0,
                 walloc=0.0):
    """fields: dict ex..bz of equal (mz, my, mx) arrays (ghosts included)"""
    shape = np.shape(fields["ex"])
    if len(shape) != 3:
        raise ValueError("fields must be (mz, my, mx) arrays")
    mz, my, mx = shape
    parts = [np.array([mx, my, mz], "<i4").tobytes()]
    for k in FIELD_NAMES:
        a = np.ascontiguousarray(fields[k], "<f4")
        if a.shape != shape:
            raise ValueError(f"{k} has shape {a.shape}, ex has {shape}")
        parts.append(a.tobytes())
    parts += [np.array([dseed], "<f8").tobytes(), np.array([lap], "<i4").tobytes(),
              np.array([xinject, xinject2, xinject3], "<f8").tobytes(), np.array([leftwall], "<f4").tobytes(),
              np.array([walloc], "<f8").tobytes()]
    write_record(path, parts)


def read_particles(path):
    """dict with ions, lecs, maxptl, maxhlf, totalpartnum (the counts) and
    "species": (ion arrays, electron arrays), each a dict of x, y, z, u, v,
    w, ch, ind, proc, splitlev"""
    c = _Cursor(read_record(path), path)
    ions, lecs, maxptl, maxhlf, total = (int(v) for v in c.take("<i4", 5))
    if ions < 0 or lecs < 0:
        raise ValueError(f"{path}: particle counts {ions}, {lecs}")
    out = dict(ions=ions, lecs=lecs, maxptl=maxptl, maxhlf=maxhlf, totalpartnum=total)
    sp = ({}, {})
    for names, dt in ((PART_REALS, "<f4"), (PART_INTS, "<i4")):
        for k in names:
            sp[0][k] = c.take(dt, ions)
            sp[1][k] = c.take(dt, lecs)
    c.done()
    out["species"] = sp
    return out


def write_particles(path, ions, lecs, maxptl=None, maxhlf=None, totalpartnum=None):
    """ions, lecs: dicts of equal-length arrays x..ch (ind, proc, splitlev
    default to 0)"""
    ni, ne = len(ions["x"]), len(lecs["x"])
    maxhlf = max(ni, ne) if maxhlf is None else maxhlf
    maxptl = 2 * maxhlf if maxptl is None else maxptl
    totalpartnum = ni + ne if totalpartnum is None else totalpartnum
    parts = [np.array([ni, ne, maxptl, maxhlf, totalpartnum], "<i4").tobytes()]
    for names, dt in ((PART_REALS, "<f4"), (PART_INTS, "<i4")):
        for k in names:
            for d, n in ((ions, ni), (lecs, ne)):
                a = np.asarray(d[k], dt) if k in d else np.zeros(n, dt)
                if a.shape != (n,):
                    raise ValueError(f"{k} has shape {a.shape}, expected ({n},)")
                parts.append(a.tobytes())
    write_record(path, parts)
