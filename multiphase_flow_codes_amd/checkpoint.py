"""Checkpoint files of the reference, read and written from a Lattice.

The reference's restart files (one rank, myid = 0) are Fortran sequential
unformatted files:

  endrunflow2D8x8S1.<istep:06d>.<myid:03d>   (savecntdflow / loadcntdflow,
      BOX/saveload.f90:186-214, 287-310)
      record 1: nsteps, istat, imovie                  int32 x 3
      record 2: f(0:18, lx, ly, lz)                    float64, population fastest
  endrunpart2D.<istep:06d>.<myid:03d>        (savecntdpart / loadcntdpart,
      BOX/saveload.f90:582-618, 621-665)
      record 1: nps, ipglb(msize)                      int32
      record 2: ibnodes0(lx,ly,lz), isnodes0(lx,ly,lz) int32
      record 3: yp(3,msize), thetap(3,msize)           float64
      record 4: forcepp(3,msize), torqpp(3,msize)      float64
      record 5: dwdt(3,msize), domgdt(3,msize)         float64
      record 6: ypglb(3,npart), wp(3,npart), omgp(3,npart)  float64 (myid 0)
  force.<istep:06d>                          (input_outputf,
      BOX/saveload.f90:218-274)
      records: iseedf; ivf(32); iyf; b1r(6,5,5), b2r, b3r   int32 / float64

msize = int(5 npart / nproc) (BOX/para.f90:248) sizes the rank-local arrays;
on one rank local slot ip holds particle ipglb(ip) = ip.  Every record is
framed by 4-byte little-endian length markers; a record longer than
2,147,483,639 bytes is written as subrecords whose leading marker is negated
when another subrecord follows and whose trailing marker is negated when
one precedes (the convention of the Intel and GNU Fortran runtimes for
records over 2 GB).

A restart continues as the reference's does: f, the particles' positions,
velocities, spins and the Newton-Euler history, the forcing process state;
the node maps are rebuilt by beads_links and the rho/u arrays by macrovar
(BOX/main.f90:123-140).  ibnodes0/isnodes0 are written for the reference and
read back for inspection: beads_move replaces them before anything reads
them.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAX_SUBRECORD = 2_147_483_639


def _stamp(istep: int) -> str:
    return f"{int(istep):06d}"


def flow_filename(istep: int, myid: int = 0) -> str:
    return f"endrunflow2D8x8S1.{_stamp(istep)}.{int(myid):03d}"


def part_filename(istep: int, myid: int = 0) -> str:
    return f"endrunpart2D.{_stamp(istep)}.{int(myid):03d}"


def force_filename(istep: int) -> str:
    return f"force.{_stamp(istep)}"


# ---- sequential unformatted records ---------------------------------------------
def write_record(fh, arrays, max_sub: int = MAX_SUBRECORD):
    """One record holding the arrays' bytes back to back (native order)."""
    views = [memoryview(np.ascontiguousarray(a)).cast("B") for a in arrays]
    total = sum(v.nbytes for v in views)
    if total == 0:
        fh.write(struct.pack("<ii", 0, 0))
        return
    # walk the concatenated bytes subrecord by subrecord
    nsub = (total + max_sub - 1) // max_sub
    vi, off = 0, 0
    for s in range(nsub):
        n = min(max_sub, total - s * max_sub)
        fh.write(struct.pack("<i", -n if s < nsub - 1 else n))
        left = n
        while left:
            v = views[vi]
            k = min(left, v.nbytes - off)
            fh.write(v[off:off + k])
            left -= k
            off += k
            if off == v.nbytes:
                vi, off = vi + 1, 0
        fh.write(struct.pack("<i", -n if s > 0 else n))


def read_record(fh, expect: int | None = None) -> np.ndarray:
    """The payload of the next record as one uint8 array, each subrecord read
    straight into it (no second copy).  `expect`, when given, is the payload
    length the caller needs (preallocated; a different length is an error).

    Two framings are read: subrecords of at most 2^31 - 9 bytes whose markers
    carry the sign of continuation (the GNU and Intel runtimes' records over
    2 GB, as written here), and a record of 2^31 bytes or more framed by a
    single pair of markers holding its length as an unsigned 32-bit value
    (the runtime of amdflang).  A negative first marker is either; the
    trailing marker tells them apart: +|n| after |n| bytes for a first
    subrecord, the same negative value after 2^32 + n bytes otherwise."""
    buf = np.empty(expect, dtype=np.uint8) if expect is not None else None
    parts = []
    pos = 0
    first = True
    while True:
        head = fh.read(4)
        if len(head) < 4:
            raise EOFError("end of file inside a record" if not first else "end of file")
        (h,) = struct.unpack("<i", head)
        n, more = abs(h), h < 0
        if h < 0 and first:
            start = fh.tell()
            fh.seek(start + n)
            t = fh.read(4)
            if len(t) < 4 or struct.unpack("<i", t)[0] != n:
                # one record with an unsigned 32-bit length (amdflang)
                fh.seek(start + h + (1 << 32))
                t = fh.read(4)
                if len(t) < 4 or struct.unpack("<i", t)[0] != h:
                    raise ValueError("corrupt record framing")
                n, more = h + (1 << 32), False
            fh.seek(start)
        if buf is not None:
            if pos + n > buf.size:
                raise ValueError(f"record longer than the expected {expect} bytes")
            got = fh.readinto(memoryview(buf)[pos:pos + n])
        else:
            data = fh.read(n)
            got = len(data)
            parts.append(data)
        pos += got
        t = fh.read(4)
        if got != n or len(t) < 4:
            raise ValueError("corrupt record framing")
        (tail,) = struct.unpack("<i", t)
        if n < (1 << 31):
            if abs(tail) != n:
                raise ValueError("corrupt record framing")
            if (tail < 0) == first:
                raise ValueError("subrecord continuation markers do not match")
        first = False
        if not more:
            break
    if buf is None:
        return np.frombuffer(b"".join(parts), dtype=np.uint8)
    if pos != buf.size:
        raise ValueError(f"record holds {pos} bytes, expected {expect}")
    return buf


def read_records(path):
    out = []
    with open(path, "rb") as fh:
        while True:
            try:
                out.append(read_record(fh))
            except EOFError as e:
                if str(e) == "end of file":
                    return out
                raise


def _split(buf, spec):
    """Cut a record into arrays (views of it): spec = [(dtype, count), ...]."""
    buf = np.asarray(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
    out, off = [], 0
    for dt, n in spec:
        a = np.frombuffer(buf, dtype=dt, count=n, offset=off)
        out.append(a)
        off += a.nbytes
    if off != buf.size:
        raise ValueError(f"record holds {buf.size} bytes, expected {off}")
    return out


# ---- the three files -------------------------------------------------------------------
def save_flow(path, f, nsteps: int, istat: int = 0, imovie: int = 0, max_sub: int = MAX_SUBRECORD):
    """f: (nz, ny, nx, 19) float64 (Lattice.download_f), i.e. f(0:18,lx,ly,lz)."""
    with open(path, "wb") as fh:
        write_record(fh, [np.array([nsteps, istat, imovie], dtype="<i4")], max_sub)
        write_record(fh, [np.ascontiguousarray(f, dtype="<f8")], max_sub)


def load_flow(path, shape):
    """(nsteps, istat, imovie), f as (nz, ny, nx, 19).  The population
    record is read straight into the returned array (one copy in memory)."""
    nz, ny, nx = shape
    with open(path, "rb") as fh:
        (hdr,) = _split(read_record(fh, 12), [("<i4", 3)])
        (f,) = _split(read_record(fh, 8 * 19 * nx * ny * nz), [("<f8", 19 * nx * ny * nz)])
        if fh.read(1):
            raise ValueError(f"{path}: more than 2 records")
    return tuple(int(v) for v in hdr), f.reshape(nz, ny, nx, 19)


PART_FIELDS = ("yp", "thetap", "forcepp", "torqpp", "dwdt", "domgdt")


def save_part(path, npart, shape, st: dict, msize: int | None = None, max_sub: int = MAX_SUBRECORD):
    """st: ypglb, wp, omgp and the local arrays of PART_FIELDS as (npart, 3)
    (slot = particle id - 1), ibnodes0 / isnodes0 as (nz, ny, nx) int32."""
    msize = 5 * npart if msize is None else msize
    if msize < npart:
        raise ValueError("msize must hold every particle on one rank")

    def local(a):
        out = np.zeros((msize, 3), dtype="<f8")
        out[:npart] = np.asarray(a, dtype=np.float64).reshape(npart, 3)
        return out

    ipglb = np.zeros(msize, dtype="<i4")
    ipglb[:npart] = np.arange(1, npart + 1)
    with open(path, "wb") as fh:
        write_record(fh, [np.array([npart], dtype="<i4"), ipglb], max_sub)
        write_record(fh, [np.asarray(st["ibnodes0"], dtype="<i4").reshape(shape),
                          np.asarray(st["isnodes0"], dtype="<i4").reshape(shape)], max_sub)
        write_record(fh, [local(st["yp"]), local(st["thetap"])], max_sub)
        write_record(fh, [local(st["forcepp"]), local(st["torqpp"])], max_sub)
        write_record(fh, [local(st["dwdt"]), local(st["domgdt"])], max_sub)
        write_record(fh, [np.asarray(st[k], dtype="<f8").reshape(npart, 3) for k in ("ypglb", "wp", "omgp")],
                     max_sub)


def load_part(path, npart, shape):
    """Inverse of save_part (also for files of the reference): the local
    arrays are put back in particle-id order through ipglb."""
    ncell = int(np.prod(shape))
    recs = read_records(path)
    if len(recs) != 6:
        raise ValueError(f"{path}: {len(recs)} records, expected 6")
    msize = (len(recs[0]) - 4) // 4
    nps, ipglb = _split(recs[0], [("<i4", 1), ("<i4", msize)])
    nps = int(nps[0])
    ib0, is0 = _split(recs[1], [("<i4", ncell), ("<i4", ncell)])
    st = {"nps": nps, "msize": msize, "ibnodes0": ib0.reshape(shape), "isnodes0": is0.reshape(shape)}
    ids = ipglb[:nps] - 1
    if nps != npart or np.any(ids < 0) or np.any(ids >= npart) or len(set(ids.tolist())) != nps:
        raise ValueError("the file does not hold every particle of one rank")
    for rec, names in zip(recs[2:5], (("yp", "thetap"), ("forcepp", "torqpp"), ("dwdt", "domgdt"))):
        a, b = _split(rec, [("<f8", 3 * msize), ("<f8", 3 * msize)])
        for name, arr in ((names[0], a), (names[1], b)):
            out = np.zeros((npart, 3))
            out[ids] = arr.reshape(msize, 3)[:nps]
            st[name] = out
    yg, wp, om = _split(recs[5], [("<f8", 3 * npart)] * 3)
    st["ypglb"], st["wp"], st["omgp"] = yg.reshape(npart, 3), wp.reshape(npart, 3), om.reshape(npart, 3)
    return st


def save_force(path, iseedf, ivf, iyf, b, max_sub: int = MAX_SUBRECORD):
    """b: (b1r, b2r, b3r), each the 150 values of b?r(6,5,5) in Fortran order."""
    with open(path, "wb") as fh:
        write_record(fh, [np.array([iseedf], dtype="<i4")], max_sub)
        write_record(fh, [np.asarray(ivf, dtype="<i4").reshape(32)], max_sub)
        write_record(fh, [np.array([iyf], dtype="<i4")], max_sub)
        write_record(fh, [np.asarray(x, dtype="<f8").reshape(150) for x in b], max_sub)


def load_force(path):
    recs = read_records(path)
    if len(recs) != 4:
        raise ValueError(f"{path}: {len(recs)} records, expected 4")
    (iseedf,) = _split(recs[0], [("<i4", 1)])
    (ivf,) = _split(recs[1], [("<i4", 32)])
    (iyf,) = _split(recs[2], [("<i4", 1)])
    b = _split(recs[3], [("<f8", 150)] * 3)
    return {"iseedf": int(iseedf[0]), "ivf": ivf, "iyf": int(iyf[0]), "b": b}


# ---- a Lattice's state ---------------------------------------------------------------
def save_lattice(lat, directory, istep: int, istat: int = 0, imovie: int = 0):
    """The three files of the reference at step istep (the state after that
    step's macrovar).  Returns their paths."""
    os.makedirs(directory, exist_ok=True)
    paths = {"flow": os.path.join(directory, flow_filename(istep))}
    save_flow(paths["flow"], lat.download_f(), istep, istat, imovie)
    npart = lat.prm.npart
    if npart > 0:
        st = {k: lat.prt_get(k) for k in ("ypglb", "wp", "omgp", "thetap", "forcepp", "torqpp",
                                          "dwdt", "domgdt")}
        st["yp"] = st["ypglb"]      # one rank: the local copy is the global one after the wrap
        ib0, is0 = lat.get_maps(1)
        st["ibnodes0"], st["isnodes0"] = ib0, is0
        paths["part"] = os.path.join(directory, part_filename(istep))
        save_part(paths["part"], npart, lat.shape, st)
    if lat.prm.forcing:
        fs = lat.get_forcing_state()
        paths["force"] = os.path.join(directory, force_filename(istep))
        save_force(paths["force"], fs["iseedf"], fs["ivf"], fs["iyf"], fs["b"])
    return paths


def restore_lattice(lat, directory, istep: int):
    """Load the files of step istep into a freshly created Lattice of the same
    run constants, as the reference's restart does (loadcntdflow,
    loadcntdpart, input_outputf, then beads_links and macrovar).  Returns the
    flow file's (nsteps, istat, imovie)."""
    hdr, f = load_flow(os.path.join(directory, flow_filename(istep)), lat.shape)
    lat.upload_f(f)
    npart = lat.prm.npart
    if npart > 0:
        st = load_part(os.path.join(directory, part_filename(istep)), npart, lat.shape)
        lat.prt_init(npart)
        lat.prt_set("ypglb", st["ypglb"])
        for k in ("wp", "omgp", "thetap", "forcepp", "torqpp", "dwdt", "domgdt"):
            lat.prt_set(k, st[k])
        lat.beads_links()
    if lat.prm.forcing:
        fp = os.path.join(directory, force_filename(istep))
        if os.path.exists(fp):
            fs = load_force(fp)
            lat.set_forcing_state(fs["iseedf"], fs["iyf"], fs["ivf"], fs["b"])
    lat.macrovar()
    return hdr
