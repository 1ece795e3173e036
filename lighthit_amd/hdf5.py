"""Read-only access to HDF5 files in libhdf5's default ("earliest") layout.

The G4 step files of the reference are HDF5 (``G4Shower.from_hdf5``,
``sources.py:362-393``; the reference reads them with its own small reader,
``hdf5_minimal.py``, where h5py is absent, as it is here). This module reads
that layout from the HDF5 file format specification, nothing else:

* superblock version 0 (the root group's symbol-table entry);
* old-style groups: a symbol-table message pointing to a version-1 B-tree of
  symbol-table nodes and a local heap holding the link names;
* version-1 object headers (with continuation blocks);
* dataspace (versions 1 and 2), datatype (fixed-point, floating-point,
  fixed-length string, compound of versions 1-3, array) and data-layout
  (versions 1-3: compact, contiguous, chunked through a version-1 B-tree)
  messages.

Datasets of any rank come back as numpy arrays in the file's byte order (a
compound type as a structured dtype with the file's member offsets and
size). A filter pipeline (compression, shuffle) is refused, as is anything
the layout above does not cover. Nothing is written.
"""
from __future__ import annotations

import numpy as np

__all__ = ["File", "read_dataset", "list_datasets"]

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEFINED = 0xFFFFFFFFFFFFFFFF

# message types of a version-1 object header
_DATASPACE, _DATATYPE, _LAYOUT, _FILTERS, _CONTINUATION, _SYMBOLS = 0x1, 0x3, 0x8, 0xB, 0x10, 0x11


class _Reader:
    """Little-endian fields of a byte buffer at increasing offsets."""

    def __init__(self, data, at):
        self.data, self.at = data, at

    def u(self, size):
        value = int.from_bytes(bytes(self.data[self.at:self.at + size]), "little")
        self.at += size
        return value

    def skip(self, size):
        self.at += size

    def raw(self, size):
        out = bytes(self.data[self.at:self.at + size])
        self.at += size
        return out


class File:
    """An HDF5 file mapped read-only; ``file[path]`` is the dataset at ``path``
    ("group/subgroup/name") as a numpy array."""

    def __init__(self, path):
        self.data = np.memmap(path, dtype=np.uint8, mode="r")
        if bytes(self.data[:8]) != SIGNATURE:
            raise ValueError("not an HDF5 file")
        if int(self.data[8]) != 0:
            raise ValueError(f"superblock version {int(self.data[8])} is not supported "
                             "(only the earliest layout, version 0)")
        r = _Reader(self.data, 13)
        self.offset_size, self.length_size = r.u(1), r.u(1)
        if self.offset_size not in (4, 8) or self.length_size not in (4, 8):
            raise ValueError("unsupported offset or length size")
        r.skip(1 + 2 + 2 + 4)                   # reserved, group K values, consistency flags
        self.base = r.u(self.offset_size)
        r.skip(3 * self.offset_size)            # free space, end of file, driver info
        self.root = self._entry(r)["header"]

    # ------------------------------------------------------------ structures

    def _o(self, r):
        return r.u(self.offset_size)

    def _l(self, r):
        return r.u(self.length_size)

    def _entry(self, r):
        """A symbol-table entry: link name offset and object header address."""
        name = self._o(r)
        header = self._o(r)
        r.skip(4 + 4 + 16)                      # cache type, reserved, scratch pad
        return {"name": name, "header": header}

    def _messages(self, address):
        """Messages (type, body offset, body size) of the version-1 object header at
        ``address``, following continuation blocks."""
        r = _Reader(self.data, self.base + address)
        if r.u(1) != 1:
            raise ValueError("only version-1 object headers are supported")
        r.skip(1)
        count = r.u(2)
        r.skip(4)                               # reference count
        size = r.u(4)
        blocks = [(self.base + address + 16, size)]
        out = []
        while blocks and len(out) < count:
            start, length = blocks.pop(0)
            r = _Reader(self.data, start)
            while r.at + 8 <= start + length and len(out) < count:
                kind, body = r.u(2), r.u(2)
                r.skip(4)                       # flags, reserved
                if kind == _CONTINUATION:
                    c = _Reader(self.data, r.at)
                    blocks.append((self.base + self._o(c), self._l(c)))
                out.append((kind, r.at, body))
                r.skip(body)
        return out

    def _children(self, header):
        """Names and object header addresses of a group's links."""
        table = [m for m in self._messages(header) if m[0] == _SYMBOLS]
        if not table:
            raise ValueError("not a group (no symbol-table message)")
        r = _Reader(self.data, table[0][1])
        tree, heap = self._o(r), self._o(r)
        h = _Reader(self.data, self.base + heap)
        if h.raw(4) != b"HEAP":
            raise ValueError("bad local heap")
        h.skip(4)
        self._l(h)
        self._l(h)
        names = self.base + self._o(h)
        out = {}
        for node in self._tree_children(tree, 0):
            s = _Reader(self.data, self.base + node)
            if s.raw(4) != b"SNOD":
                raise ValueError("bad symbol-table node")
            s.skip(2)
            for _ in range(s.u(2)):
                entry = self._entry(s)
                at = names + entry["name"]
                end = at
                while self.data[end] != 0:
                    end += 1
                out[bytes(self.data[at:end]).decode()] = entry["header"]
        return out

    def _tree_children(self, address, kind, rank=0):
        """Leaf children of a version-1 B-tree of ``kind`` 0 (group: addresses of
        symbol-table nodes) or 1 (chunks: (chunk size, filter mask, offsets,
        address))."""
        r = _Reader(self.data, self.base + address)
        if r.raw(4) != b"TREE":
            raise ValueError("bad B-tree node")
        if r.u(1) != kind:
            raise ValueError("unexpected B-tree node type")
        level, used = r.u(1), r.u(2)
        r.skip(2 * self.offset_size)            # siblings
        out = []
        for _ in range(used):
            if kind == 0:
                self._l(r)                      # key: heap offset of a name
                child = self._o(r)
                out.extend(self._tree_children(child, 0) if level else [child])
            else:
                size, mask = r.u(4), r.u(4)
                offsets = [r.u(8) for _ in range(rank + 1)]
                child = self._o(r)
                out.extend(self._tree_children(child, 1, rank) if level
                           else [(size, mask, offsets[:rank], child)])
        return out

    # ------------------------------------------------------------- messages

    def _dataspace(self, at):
        r = _Reader(self.data, at)
        version, rank, flags = r.u(1), r.u(1), r.u(1)
        if version == 1:
            r.skip(5)
        elif version == 2:
            r.skip(1)                           # dataspace type
        else:
            raise ValueError(f"dataspace version {version} is not supported")
        return tuple(self._l(r) for _ in range(rank))

    def _datatype(self, r):
        """numpy dtype of the datatype message at the reader's position."""
        head = r.u(1)
        kind, version = head & 0x0F, head >> 4
        bits = r.u(3)
        size = r.u(4)
        order = ">" if bits & 1 else "<"
        if kind == 0:                           # fixed point
            r.skip(4)
            if size not in (1, 2, 4, 8):
                raise ValueError(f"{size}-byte integers are not supported")
            return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}")
        if kind == 1:                           # floating point
            r.skip(12)
            if size not in (2, 4, 8):
                raise ValueError(f"{size}-byte floats are not supported")
            return np.dtype(f"{order}f{size}")
        if kind == 3:                           # fixed-length string
            return np.dtype(f"S{size}")
        if kind == 6:                           # compound
            members = bits & 0xFFFF
            names, formats, offsets = [], [], []
            for _ in range(members):
                start = r.at
                end = start
                while self.data[end] != 0:
                    end += 1
                names.append(bytes(self.data[start:end]).decode())
                if version < 3:
                    r.at = start + ((end - start) // 8 + 1) * 8
                    offsets.append(r.u(4))
                else:
                    r.at = end + 1
                    width = 1 if size < 256 else 2 if size < 65536 else 3 if size < 2 ** 24 else 4
                    offsets.append(r.u(width))
                shape = ()
                if version == 1:
                    rank = r.u(1)
                    r.skip(3 + 4 + 4)
                    dims = [r.u(4) for _ in range(4)]
                    shape = tuple(dims[:rank])
                member = self._datatype(r)
                formats.append(np.dtype((member, shape)) if shape else member)
            return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                             "itemsize": size})
        if kind == 10:                          # array
            rank = r.u(1)
            if version < 3:
                r.skip(3)
            dims = tuple(r.u(4) for _ in range(rank))
            if version < 3:
                r.skip(4 * rank)                # permutation
            return np.dtype((self._datatype(r), dims))
        raise ValueError(f"datatype class {kind} is not supported")

    def _layout(self, at, rank):
        r = _Reader(self.data, at)
        version = r.u(1)
        if version in (1, 2):
            dims_count = r.u(1)
            kind = r.u(1)
            r.skip(5)
            address = self._o(r) if kind != 0 else None
            dims = [r.u(4) for _ in range(dims_count)]
            if kind == 0:
                size = r.u(4)
                return ("compact", r.at, size)
            if kind == 1:
                return ("contiguous", address, None)
            return ("chunked", address, dims[:rank])
        if version != 3:
            raise ValueError(f"data layout version {version} is not supported")
        kind = r.u(1)
        if kind == 0:
            size = r.u(2)
            return ("compact", r.at, size)
        if kind == 1:
            address = self._o(r)
            self._l(r)
            return ("contiguous", address, None)
        if kind == 2:
            dims_count = r.u(1)
            address = self._o(r)
            dims = [r.u(4) for _ in range(dims_count)]
            return ("chunked", address, dims[:rank])
        raise ValueError(f"layout class {kind} is not supported")

    # ------------------------------------------------------------ interface

    def _object(self, path):
        header = self.root
        for part in [p for p in str(path).split("/") if p]:
            children = self._children(header)
            if part not in children:
                raise KeyError(f"{path}: no object {part!r}")
            header = children[part]
        return header

    def keys(self, group="/"):
        return sorted(self._children(self._object(group)))

    def __getitem__(self, path):
        messages = self._messages(self._object(path))
        kinds = {m[0]: m for m in messages}
        if _FILTERS in kinds:
            raise ValueError(f"{path}: filtered (compressed) datasets are not supported")
        if _DATASPACE not in kinds or _DATATYPE not in kinds or _LAYOUT not in kinds:
            raise ValueError(f"{path}: not a dataset")
        shape = self._dataspace(kinds[_DATASPACE][1])
        dtype = self._datatype(_Reader(self.data, kinds[_DATATYPE][1]))
        layout, where, chunk = self._layout(kinds[_LAYOUT][1], len(shape))
        count = int(np.prod(shape, dtype=np.int64)) if shape else 1
        if layout == "compact":
            raw = np.frombuffer(bytes(self.data[where:where + count * dtype.itemsize]), dtype)
            return raw.reshape(shape).copy()
        if layout == "contiguous":
            if where == UNDEFINED:
                return np.zeros(shape, dtype)
            start = self.base + where
            raw = np.frombuffer(bytes(self.data[start:start + count * dtype.itemsize]), dtype)
            return raw.reshape(shape).copy()
        out = np.zeros(shape, dtype)
        if where == UNDEFINED:
            return out
        for size, mask, offsets, address in self._tree_children(where, 1, len(shape)):
            if mask:
                raise ValueError(f"{path}: filtered chunks are not supported")
            block = np.frombuffer(bytes(self.data[self.base + address:
                                                  self.base + address + size]), dtype)
            block = block[:int(np.prod(chunk))].reshape(chunk)
            region = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offsets, chunk, shape))
            out[region] = block[tuple(slice(0, r.stop - r.start) for r in region)]
        return out


def read_dataset(path, name):
    """The dataset ``name`` of the HDF5 file at ``path`` as a numpy array."""
    return File(path)[name]


def list_datasets(path, group="/"):
    """Names of the links in ``group``."""
    return File(path).keys(group)
