"""Trajectory utilities of the reference's orbc-util (src/orbc-util.cpp:32-300).

    python -m openrbc_amd.orbc_util convert in.orbc out.lammpstrj
    python -m openrbc_amd.orbc_util convert in.orbc 'out.%d.namdbin'
    python -m openrbc_amd.orbc_util extract in.orbc out.orbc 100 200 ...

The frame layout is the reference's (src/trajectory.h:26-150): 8-byte
section titles, FRAMEBEG int32 step, NATOM uint64, IDENTITY (int32 tag,
int32 type) per atom, FP32 triples for POSITION / VELOCITY / ROTATION /
FORCE, int32 per atom for VORONOI, FRAMEEND.  The outputs match the
reference tool byte for byte, including its bookkeeping across frames: the
arrays persist from frame to frame (a section missing from a frame keeps the
previous frame's values), the NAMD writer keeps the largest atom count seen,
and extract writes the sections of every field met so far.
"""
from __future__ import annotations

import struct
import sys

import numpy as np

_TRIPLES = ("POSITION", "VELOCITY", "ROTATION", "FORCE")
# DumpField bits, src/runtime_parameter.h:27-33
_BIT = {"POSITION": 1, "ROTATION": 2, "VORONOI": 4, "VELOCITY": 8, "FORCE": 16}


class _Frames:
    """Sequential reader of the section stream (src/orbc-util.cpp:55-100)."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        self.pos = 0

    def title(self):
        if self.pos + 8 > len(self.buf):
            return None
        t = self.buf[self.pos:self.pos + 8].split(b"\0", 1)[0].decode("latin-1")
        self.pos += 8
        return t

    def take(self, dtype, count):
        dt = np.dtype(dtype)
        nbytes = dt.itemsize * count
        if self.pos + nbytes > len(self.buf):
            raise ValueError("truncated trajectory")
        a = np.frombuffer(self.buf, dtype=dt, count=count, offset=self.pos).copy()
        self.pos += nbytes
        return a


def _resize(a, n, fill=0):
    if len(a) >= n:
        return a[:n].copy() if len(a) > n else a
    out = np.full((n,) + a.shape[1:], fill, dtype=a.dtype)
    out[:len(a)] = a
    return out


def _fmt(v) -> str:
    """A float as C++ iostreams print it by default (6 significant digits)."""
    return "%g" % float(v)


class _State:
    """The arrays the reference tool keeps between frames."""

    def __init__(self):
        self.n = 0
        self.tag = np.zeros(0, np.int32)
        self.type = np.zeros(0, np.int32)
        self.aff = np.zeros(0, np.int32)
        self.fields = {k: np.zeros((0, 3), np.float32) for k in _TRIPLES}
        self.step = 0
        # RTParameter's default dump field, src/runtime_parameter.h:69
        self.dump_field = _BIT["POSITION"] | _BIT["ROTATION"] | _BIT["VORONOI"]

    def resize(self, n):
        self.n = n
        self.tag = _resize(self.tag, n)
        self.type = _resize(self.type, n)
        self.aff = _resize(self.aff, n)
        for k in _TRIPLES:
            self.fields[k] = _resize(self.fields[k], n)


def _read_frame(r: _Frames, st: _State, natom_hook=None, identity_hook=None) -> bool:
    """One frame into st; False at the end of the file."""
    while True:
        t = r.title()
        if t is None:
            return False
        if t == "FRAMEBEG":
            st.step = int(r.take("<i4", 1)[0])
        elif t == "FRAMEEND":
            return True
        elif t == "NATOM":
            n = int(r.take("<u8", 1)[0])
            if natom_hook:
                natom_hook(n)
            else:
                st.resize(n)
        elif t == "IDENTITY":
            m = identity_hook() if identity_hook else st.n
            ids = r.take("<i4", 2 * m).reshape(m, 2)
            st.tag[:m] = ids[:, 0]
            st.type[:m] = ids[:, 1]
            if identity_hook:
                st.tag[m:] = st.n + 1
        elif t in _TRIPLES:
            st.dump_field |= _BIT[t]
            m = identity_hook() if identity_hook else st.n
            st.fields[t][:m] = r.take("<f4", 3 * m).reshape(m, 3)
            if identity_hook and t == "POSITION":
                st.fields[t][m:] = 0.0
        elif t == "VORONOI":
            st.dump_field |= _BIT[t]
            m = identity_hook() if identity_hook else st.n
            st.aff[:m] = r.take("<i4", m)
        else:
            sys.stdout.write("Unknown section %s in file\n" % t)


def _by_tag(st: _State):
    return np.argsort(st.tag, kind="stable")


def to_lammpstrj(src: str, dst: str) -> int:
    """orbc2lammpstrj, src/orbc-util.cpp:52-140: every frame as a LAMMPS dump
    of id, type, unwrapped position and Voronoi cell (the column the tool
    labels `vx`), atoms in tag order, the default box of +-1000.  Returns the
    number of frames written."""
    r, st = _Frames(src), _State()
    frames = 0
    with open(dst, "w", newline="\n") as out:
        while _read_frame(r, st):
            x = st.fields["POSITION"]
            lines = ["ITEM: TIMESTEP", str(st.step), "ITEM: NUMBER OF ATOMS", str(st.n),
                     "ITEM: BOX BOUNDS ff ff ff", "-1000 1000", "-1000 1000", "-1000 1000",
                     "ITEM: ATOMS id type xu yu zu vx"]
            for i in _by_tag(st):
                lines.append(f"{st.tag[i]} {st.type[i]} {_fmt(x[i, 0])} {_fmt(x[i, 1])} {_fmt(x[i, 2])} {st.aff[i]} ")
            out.write("\n".join(lines) + "\n")
            frames += 1
    return frames


def to_namdbin(src: str, dst_pattern: str) -> list:
    """orbc2namdbin, src/orbc-util.cpp:142-232: one NAMD binary coordinate
    file per frame (dst_pattern % step): uint32 atom count, then FP64 x y z
    per atom in tag order.  The atom count is the largest met so far; atoms
    beyond the frame's own count are written at the origin.  Returns the
    files written."""
    r, st = _Frames(src), _State()
    eff = [0]

    def natom(n):
        eff[0] = n
        if st.n < n:
            st.resize(n)

    written = []
    while _read_frame(r, st, natom_hook=natom, identity_hook=lambda: eff[0]):
        name = dst_pattern % st.step
        x = st.fields["POSITION"].astype(np.float64)
        with open(name, "wb") as out:
            out.write(struct.pack("<I", st.n))
            out.write(np.ascontiguousarray(x[_by_tag(st)]).tobytes())
        written.append(name)
    return written


def extract(src: str, dst: str, steps) -> list:
    """extract, src/orbc-util.cpp:237-292 over read_frame,
    src/trajectory.h:108-150: the frames whose step is listed, rewritten with
    POSITION and ROTATION (the default dump field) plus every other triple
    met so far in the file (the dump-field bits accumulate), Voronoi cells not
    written.  Returns
    the steps written."""
    want = set(int(s) for s in steps)
    r, st = _Frames(src), _State()
    done = []
    with open(dst, "wb") as out:
        def title(t):
            out.write(t.encode("latin-1").ljust(8, b"\0"))
        while _read_frame(r, st):
            if st.step not in want:
                continue
            title("FRAMEBEG"); out.write(struct.pack("<i", st.step))
            title("NATOM"); out.write(struct.pack("<Q", st.n))
            title("IDENTITY")
            out.write(np.stack([st.tag, st.type], 1).astype("<i4").tobytes())
            for t in ("POSITION", "VELOCITY", "ROTATION", "FORCE"):
                if st.dump_field & _BIT[t]:
                    title(t)
                    out.write(st.fields[t].astype("<f4").tobytes())
            title("FRAMEEND")
            done.append(st.step)
    return done


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("convert", "extract"):
        sys.stdout.write("List of commands:\nconvert\nextract\nhelp\n")
        return 0
    if argv[0] == "convert":
        if len(argv) < 3:
            sys.stdout.write("Usage: orbc_util convert input_filename output_filename\n")
            return 255
        src, dst = argv[1], argv[2]
        if src.rsplit(".", 1)[-1] == "orbc" and dst.rsplit(".", 1)[-1] == "lammpstrj":
            to_lammpstrj(src, dst)
        if src.rsplit(".", 1)[-1] == "orbc" and dst.rsplit(".", 1)[-1] == "namdbin":
            to_namdbin(src, dst)
        return 0
    if len(argv) < 4:
        sys.stdout.write("Usage: orbc_util extract input_filename output_filename frame_ids...\n")
        return 255
    extract(argv[1], argv[2], argv[3:])
    return 0


if __name__ == "__main__":
    sys.exit(main())
