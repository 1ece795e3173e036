"""Writes g4_steps.h5: a small G4 step file in libhdf5's earliest layout.

Two events (``event_0/tracks``, ``event_1/tracks``), each a one-dimensional
contiguous compound dataset with the columns of the G4 step files (``uid``
int64; ``x_m``, ``y_m``, ``z_m``, ``t_ns``, ``beta``, ``step_length_m``
float64): a few particles per event, each a run of consecutive rows along a
slightly bent path, with zero-length steps, zero true paths and steps below
the Cherenkov threshold mixed in. The rows are stored beside it in
g4_steps_rows.npz.

h5py is absent here, so the file is written by the small writer below,
following the HDF5 file format specification (superblock version 0,
symbol-table groups with a version-1 B-tree and a local heap, version-1
object headers, compound datatype version 1, contiguous layout version 3).
The reader it feeds (lighthit_amd.hdf5) is pinned separately on files that
libhdf5 wrote. Run from the repository root:
``python tests/golden/make_g4_fixture.py``.
"""
import struct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
UNDEF = 0xFFFFFFFFFFFFFFFF
COLUMNS = [("uid", "<i8"), ("x_m", "<f8"), ("y_m", "<f8"), ("z_m", "<f8"), ("t_ns", "<f8"),
           ("beta", "<f8"), ("step_length_m", "<f8")]


def pad8(data):
    return data + b"\0" * (-len(data) % 8)


class Writer:
    def __init__(self):
        self.buf = bytearray(96)            # the superblock, written last

    def put(self, data):
        at = len(self.buf)
        self.buf += pad8(bytes(data))
        return at

    def header(self, messages):
        """A version-1 object header with (type, body) messages."""
        body = b"".join(struct.pack("<HHB3x", kind, len(pad8(data)), 0) + pad8(data)
                        for kind, data in messages)
        return self.put(struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body)

    def group(self, children):
        """A group of ``children`` {name: object header address}; returns the
        addresses of its object header, B-tree and local heap."""
        names = sorted(children)
        heap_data, offsets = bytearray(8), {}
        for name in names:
            offsets[name] = len(heap_data)
            heap_data += pad8(name.encode() + b"\0")
        data_at = self.put(heap_data)
        heap = self.put(b"HEAP" + bytes([0, 0, 0, 0])
                        + struct.pack("<QQQ", len(heap_data), UNDEF, data_at))
        entries = b"".join(struct.pack("<QQII16x", offsets[n], children[n], 0, 0) for n in names)
        node = self.put(b"SNOD" + struct.pack("<BBH", 1, 0, len(names)) + entries)
        last = offsets[names[-1]] if names else 0
        tree = self.put(b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
                        + struct.pack("<QQQ", 0, node, last))
        return self.header([(0x11, struct.pack("<QQ", tree, heap))]), tree, heap

    def dataset(self, rows):
        data_at = self.put(rows.tobytes())
        space = struct.pack("<BBBB4xQ", 1, 1, 0, 0, len(rows))
        members = b""
        for name, kind in COLUMNS:
            offset = rows.dtype.fields[name][1]
            entry = pad8(name.encode() + b"\0") + struct.pack("<IB3x4x4x4I", offset, 0, 0, 0,
                                                                0, 0)
            if kind == "<i8":        # fixed point, little-endian, signed
                member = struct.pack("<B3BI", 0x10, 0x08, 0, 0, 8) + struct.pack("<HH", 0, 64)
            else:                    # IEEE double, little-endian
                member = (struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8)
                          + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023))
            members += entry + member
        count = len(COLUMNS)
        kind = struct.pack("<B3BI", 0x16, count & 0xFF, count >> 8, 0, rows.dtype.itemsize)
        layout = struct.pack("<BBQQ", 3, 1, data_at, rows.nbytes)
        return self.header([(0x1, space), (0x3, kind + members), (0x8, layout)])

    def finish(self, root, tree, heap, path):
        sb = (b"\x89HDF\r\n\x1a\n" + bytes([0, 0, 0, 0, 0, 8, 8, 0])
              + struct.pack("<HHI", 4, 16, 0)
              + struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
              + struct.pack("<QQII", 0, root, 1, 0) + struct.pack("<QQ", tree, heap))
        assert len(sb) == 96
        self.buf[:96] = sb
        Path(path).write_bytes(bytes(self.buf))


def event_rows(rng, particles):
    out = []
    for uid in range(particles):
        steps = int(rng.integers(4, 12))
        start = rng.normal(scale=2.0, size=3)
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        beta = float(rng.uniform(0.6, 0.999))
        t = float(rng.uniform(0.0, 3.0))
        position = start
        for k in range(steps):
            row = np.zeros(1, np.dtype(COLUMNS))
            step = float(rng.uniform(0.002, 0.05))
            if k == 3:
                step = 0.0                                   # zero-length chord
            row["uid"] = 100 + uid
            row["x_m"], row["y_m"], row["z_m"] = position
            row["t_ns"] = t
            row["beta"] = beta
            row["step_length_m"] = 0.0 if k == 5 else step * (1 + rng.uniform(0, 0.02))
            out.append(row)
            direction = direction + rng.normal(scale=0.05, size=3)
            direction /= np.linalg.norm(direction)
            position = position + step * direction
            t += step / (beta * 0.299792458)
            beta = max(0.5, beta - float(rng.uniform(0.0, 0.05)))   # crosses the threshold
    return np.concatenate(out)


def main():
    rng = np.random.default_rng(41)
    events = [event_rows(rng, 3), event_rows(rng, 5)]
    w = Writer()
    groups = {}
    for k, rows in enumerate(events):
        dataset = w.dataset(rows)
        groups[f"event_{k}"] = w.group({"tracks": dataset})[0]
    root, tree, heap = w.group(groups)
    w.finish(root, tree, heap, HERE / "g4_steps.h5")
    np.savez(HERE / "g4_steps_rows.npz", **{f"event_{k}": r for k, r in enumerate(events)})
    print("g4_steps.h5:", [len(r) for r in events], "rows")


if __name__ == "__main__":
    main()
