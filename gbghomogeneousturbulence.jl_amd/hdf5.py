"""A self-contained HDF5 writer and reader for the checkpoint files.

The writer lays a file out the way libhdf5 does under its default ("earliest") format bounds,
which is also what HDF5.jl produces unless told otherwise:

- superblock version 0, offsets and lengths of 8 bytes, no user block;
- version-1 object headers, every message padded to 8 bytes;
- groups as symbol tables: a version-1 B-tree (group node K 16) over symbol-table nodes
  (leaf K 4, so eight entries each) whose link names live in a local heap;
- contiguous datasets, each with a version-1 dataspace, its datatype, a version-2 fill-value
  message and a version-3 layout message;
- attributes in the object header (attribute message version 1).

Value mapping (the one HDF5.jl uses where it has one): int -> 64-bit signed integer, float ->
IEEE binary64, bool -> 8-bit signed integer (0 or 1), str -> fixed-length UTF-8 string,
a sequence of numbers -> a one-dimensional array, complex arrays -> a compound of two floats
named ``r`` and ``i``.  Arrays keep numpy's C order: the first axis varies slowest, so a reader
in a column-major language (Julia, Fortran) sees the axes reversed, as for any file h5py writes.

The reader takes that layout, the one libhdf5 writes under its "latest" bounds (superblock
version 2 or 3, version-2 object headers, groups as link messages in the header) and files with
a user block; datasets may be contiguous or compact.  Chunked storage, shared or committed
datatypes, dense (fractal-heap) link and attribute storage and variable-length data are refused
by name.  Nothing in a file is executed: every value is decoded into numpy arrays, str, int and
float.

Format: "HDF5 File Format Specification Version 3.0", The HDF Group.
"""
from __future__ import annotations

import os
import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF        # the undefined address (all bits set)
_HEAP_FREE_NULL = 1                # a local heap's "no free block" offset, as libhdf5 writes it
_LEAF_K = 4                        # symbol-table node: up to 2K = 8 entries
_NODE_K = 16                       # group B-tree node: up to 2K = 32 children
_ENTRY = 40                        # a symbol-table entry with 8-byte offsets and lengths
_SNOD_SIZE = 8 + 2 * _LEAF_K * _ENTRY
_TREE_SIZE = 24 + 2 * _NODE_K * 8 + (2 * _NODE_K + 1) * 8

# object header message types
_NIL, _DATASPACE, _LINK_INFO, _DATATYPE, _FILL_OLD, _FILL, _LINK, _LAYOUT = 0, 1, 2, 3, 4, 5, 6, 8
_ATTRIBUTE, _CONTINUATION, _SYMBOL_TABLE, _MTIME, _ATTR_INFO = 12, 16, 17, 18, 21


class FormatError(ValueError):
    """The bytes are not an HDF5 file this reader understands."""


def _pad8(n: int) -> int:
    return (n + 7) & ~7


# ============================================================================================
# In-memory tree shared by the writer and the reader


class Group:
    """An HDF5 group: ``attrs`` (name -> value) and ``members`` (name -> Group or Dataset)."""

    def __init__(self):
        self.attrs: dict = {}
        self.members: dict = {}


class Dataset:
    """An HDF5 dataset: its array and its attributes."""

    def __init__(self, data: np.ndarray, attrs: dict | None = None):
        self.data = data
        self.attrs = attrs or {}


# ============================================================================================
# Datatype, dataspace and value encoding


def _dtype_message(dt: np.dtype) -> bytes:
    """The datatype message of a numpy dtype (class version 1, as libhdf5 writes)."""
    dt = np.dtype(dt)
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0x00
        return struct.pack("<B3BI", 0x10, bits, 0, 0, dt.itemsize) + struct.pack("<HH", 0, 8 * dt.itemsize)
    if dt.kind == "b":
        return _dtype_message(np.dtype(np.int8))
    if dt.kind == "f":
        if dt.itemsize == 8:
            sign, eloc, esize, msize, bias = 63, 52, 11, 52, 1023
        elif dt.itemsize == 4:
            sign, eloc, esize, msize, bias = 31, 23, 8, 23, 127
        else:
            raise ValueError(f"HDF5 writer: no float layout for {dt}")
        return (struct.pack("<B3BI", 0x11, 0x20, sign, 0, dt.itemsize)
                + struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esize, 0, msize, bias))
    if dt.kind == "c":
        part = np.dtype(f"<f{dt.itemsize // 2}")
        members = b""
        for name, offset in ((b"r", 0), (b"i", part.itemsize)):
            members += name.ljust(_pad8(len(name) + 1), b"\0")
            members += struct.pack("<IB3xI4x4I", offset, 0, 0, 0, 0, 0, 0)
            members += _dtype_message(part)
        return struct.pack("<BHBI", 0x16, 2, 0, dt.itemsize) + members
    if dt.kind == "S":
        # fixed-length, null-terminated, UTF-8 (padding type 0, character set 1)
        return struct.pack("<B3BI", 0x13, 0x10, 0, 0, dt.itemsize)
    raise ValueError(f"HDF5 writer: no datatype for {dt}")


def _dataspace_message(shape: tuple) -> bytes:
    """A version-1 dataspace: rank 0 is a scalar."""
    return struct.pack("<BBBxI", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", int(d)) for d in shape)


def _attribute_value(value):
    """(dtype, shape, raw bytes) of an attribute value."""
    if isinstance(value, (bool, np.bool_)):
        return np.dtype(np.int8), (), struct.pack("<b", int(bool(value)))
    if isinstance(value, (int, np.integer)):
        return np.dtype("<i8"), (), struct.pack("<q", int(value))
    if isinstance(value, (float, np.floating)):
        return np.dtype("<f8"), (), struct.pack("<d", float(value))
    if isinstance(value, str):
        raw = value.encode("utf-8")
        return np.dtype(f"S{max(len(raw), 1)}"), (), raw.ljust(max(len(raw), 1), b"\0")
    arr = np.asarray(value)
    if arr.dtype.kind == "b":
        arr = arr.astype(np.int8)
    elif arr.dtype.kind in "iu":
        arr = arr.astype("<i8")
    elif arr.dtype.kind == "f" or (arr.dtype.kind == "O" and arr.size == 0) or arr.size == 0:
        arr = arr.astype("<f8")
    elif arr.dtype.kind == "c":
        arr = arr.astype("<c16")
    else:
        raise ValueError(f"HDF5 writer: cannot store an attribute of type {type(value).__name__}")
    return arr.dtype, arr.shape, arr.tobytes(order="C")


def _message(mtype: int, body: bytes, flags: int = 0) -> bytes:
    """One version-1 object header message, padded to 8 bytes."""
    size = _pad8(len(body))
    if size > 0xFFFF:
        raise ValueError("HDF5 writer: a header message is larger than 64 KiB (an attribute too large for "
                         "the object header)")
    return struct.pack("<HHB3x", mtype, size, flags) + body.ljust(size, b"\0")


def _attribute_message(name: str, value) -> bytes:
    dt, shape, raw = _attribute_value(value)
    bname = name.encode("utf-8") + b"\0"
    dtm, dsm = _dtype_message(dt), _dataspace_message(shape)
    body = struct.pack("<BxHHH", 1, len(bname), len(dtm), len(dsm))
    body += bname.ljust(_pad8(len(bname)), b"\0") + dtm.ljust(_pad8(len(dtm)), b"\0")
    body += dsm.ljust(_pad8(len(dsm)), b"\0") + raw
    return _message(_ATTRIBUTE, body)


def _object_header(messages: list) -> bytes:
    """A version-1 object header: a 16-byte prefix, then the messages."""
    body = b"".join(messages)
    return struct.pack("<BxHII4x", 1, len(messages), 1, len(body)) + body


# ============================================================================================
# Writer


class _Writer:
    """Lays the file out front to back; raw arrays are written straight from numpy."""

    def __init__(self):
        self.end = 96                   # after the superblock
        self.pieces: list = []          # (address, bytes or ndarray)

    def put(self, obj) -> int:
        self.end = _pad8(self.end)
        addr = self.end
        self.pieces.append((addr, obj))
        self.end += obj.nbytes if isinstance(obj, np.ndarray) else len(obj)
        return addr

    def dataset(self, ds: Dataset) -> int:
        a = np.asarray(ds.data)
        if a.dtype.kind == "b":
            a = a.astype(np.int8)
        if a.dtype.kind not in "iufc":
            raise ValueError(f"HDF5 writer: cannot store a dataset of dtype {a.dtype}")
        a = np.asarray(a, dtype=a.dtype.newbyteorder("<"), order="C")
        addr = self.put(a) if a.nbytes else _UNDEF
        msgs = [_message(_DATASPACE, _dataspace_message(a.shape)),
                _message(_DATATYPE, _dtype_message(a.dtype), flags=1),
                # fill value version 2: allocation late, fill if set, the default value defined
                _message(_FILL, struct.pack("<BBBBI", 2, 2, 2, 1, 0), flags=1),
                _message(_LAYOUT, struct.pack("<BBQQ", 3, 1, addr, a.nbytes))]
        msgs += [_attribute_message(k, v) for k, v in ds.attrs.items()]
        return self.put(_object_header(msgs))

    def group(self, g: Group) -> tuple:
        """Write a group's members, then its heap, symbol-table nodes, B-tree and header.
        Returns (header address, B-tree address, heap address)."""
        entries = []
        for name in sorted(g.members, key=lambda s: s.encode("utf-8")):
            if not name or "/" in name or name == ".":
                raise ValueError(f"HDF5 writer: {name!r} is not a valid link name")
            m = g.members[name]
            if isinstance(m, Group):
                entries.append((name, *self.group(m)))
            else:
                entries.append((name, self.dataset(m if isinstance(m, Dataset) else Dataset(m)), None, None))
        if len(entries) > 2 * _LEAF_K * 2 * _NODE_K:
            raise ValueError(f"HDF5 writer: a group holds at most {4 * _LEAF_K * _NODE_K} members here")
        heap = bytearray(8)             # offset 0: the empty name
        offsets = []
        for name, *_ in entries:
            offsets.append(len(heap))
            raw = name.encode("utf-8") + b"\0"
            heap += raw.ljust(_pad8(len(raw)), b"\0")
        heap_addr = self.end = _pad8(self.end)
        self.put(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), _HEAP_FREE_NULL, heap_addr + 32))
        self.put(bytes(heap))
        chunks = [list(range(i, min(i + 2 * _LEAF_K, len(entries)))) for i in range(0, len(entries), 2 * _LEAF_K)]
        snods = []
        for chunk in chunks:
            node = bytearray(b"SNOD" + struct.pack("<BxH", 1, len(chunk)))
            for i in chunk:
                _, hdr, btree, lheap = entries[i]
                if btree is None:
                    node += struct.pack("<QQI4x16x", offsets[i], hdr, 0)
                else:
                    node += struct.pack("<QQI4xQQ", offsets[i], hdr, 1, btree, lheap)
            snods.append(self.put(bytes(node.ljust(_SNOD_SIZE, b"\0"))))
        tree = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), _UNDEF, _UNDEF))
        tree += struct.pack("<Q", 0)
        for chunk, addr in zip(chunks, snods):
            tree += struct.pack("<QQ", addr, offsets[chunk[-1]])
        btree_addr = self.put(bytes(tree.ljust(_TREE_SIZE, b"\0")))
        msgs = [_message(_SYMBOL_TABLE, struct.pack("<QQ", btree_addr, heap_addr))]
        msgs += [_attribute_message(k, v) for k, v in g.attrs.items()]
        return self.put(_object_header(msgs)), btree_addr, heap_addr


def write_file(path, root: Group) -> None:
    """Write `root` as a new HDF5 file at `path` (through a temporary file, then renamed)."""
    w = _Writer()
    hdr, btree, heap = w.group(root)
    eof = _pad8(w.end)
    sb = SIGNATURE + struct.pack("<8B", 0, 0, 0, 0, 0, 8, 8, 0) + struct.pack("<HHI", _LEAF_K, _NODE_K, 0)
    sb += struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
    sb += struct.pack("<QQI4xQQ", 0, hdr, 1, btree, heap)
    assert len(sb) == 96
    tmp = f"{os.fspath(path)}.tmp"
    with open(tmp, "wb") as f:
        f.write(sb)
        pos = 96
        for addr, obj in w.pieces:
            if addr > pos:
                f.write(b"\0" * (addr - pos))
            if isinstance(obj, np.ndarray):
                obj.tofile(f)
                pos = addr + obj.nbytes
            else:
                f.write(obj)
                pos = addr + len(obj)
        f.write(b"\0" * (eof - pos))
    os.replace(tmp, path)


# ============================================================================================
# Reader


class _Reader:
    def __init__(self, f):
        self.f = f
        self.base, self.so, self.sl = self._superblock()

    def read(self, addr: int, n: int) -> bytes:
        self.f.seek(self.base + addr)
        b = self.f.read(n)
        if len(b) != n:
            raise FormatError(f"HDF5 reader: the file ends inside a structure at address {addr}")
        return b

    def _uint(self, b: bytes, pos: int, n: int) -> int:
        return int.from_bytes(b[pos:pos + n], "little")

    def _superblock(self):
        self.f.seek(0, os.SEEK_END)
        size = self.f.tell()
        at = 0
        while True:
            if at + 8 > size:
                raise FormatError("HDF5 reader: no HDF5 signature (not an HDF5 file)")
            self.f.seek(at)
            if self.f.read(8) == SIGNATURE:
                break
            at = 512 if at == 0 else 2 * at
        self.f.seek(at)
        sb = self.f.read(256)
        version = sb[8]
        if version in (0, 1):
            so, sl = sb[13], sb[14]
            p = 24 + (4 if version == 1 else 0)
            base = self._uint(sb, p, so)
            p += 4 * so                                   # base, free space, EOF, driver info
            self.root_entry = sb[p:p + 2 * so + 24]
            self.root_header = self._uint(self.root_entry, so, so)
        elif version in (2, 3):
            so, sl = sb[9], sb[10]
            base = self._uint(sb, 12, so)
            self.root_header = self._uint(sb, 12 + 3 * so, so)
        else:
            raise FormatError(f"HDF5 reader: superblock version {version} is not known")
        if at:
            base = at                                     # after a user block, as libhdf5 takes it
        return base, so, sl

    def _undef(self, v: int, n: int) -> bool:
        return v == (1 << (8 * n)) - 1

    # --- object headers -----------------------------------------------------------------

    def messages(self, addr: int) -> list:
        """[(type, flags, body)] of the object header at `addr`, continuation chunks included."""
        head = self.read(addr, 16)
        out = []
        if head[:4] == b"OHDR":
            return self._messages_v2(addr)
        if head[0] != 1:
            raise FormatError(f"HDF5 reader: object header version {head[0]} at {addr} is not known")
        nmsgs, size = struct.unpack_from("<H", head, 2)[0], struct.unpack_from("<I", head, 8)[0]
        chunks = [(addr + 16, size)]
        while chunks and len(out) < nmsgs:
            start, length = chunks.pop(0)
            b = self.read(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsgs:
                mtype, msize, mflags = struct.unpack_from("<HHB", b, p)
                body = b[p + 8:p + 8 + msize]
                if mtype == _CONTINUATION:
                    chunks.append((self._uint(body, 0, self.so), self._uint(body, self.so, self.sl)))
                out.append((mtype, mflags, body))
                p += 8 + msize
        return out

    def _messages_v2(self, addr: int) -> list:
        head = self.read(addr, 32)
        flags = head[5]
        p = 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
        width = 1 << (flags & 3)
        size = self._uint(head, p, width)
        p += width
        out, chunks = [], [(addr + p, size, False)]
        track_order = bool(flags & 0x04)
        while chunks:
            start, length, is_cont = chunks.pop(0)
            if is_cont:
                b = self.read(start, length)
                if b[:4] != b"OCHK":
                    raise FormatError(f"HDF5 reader: missing continuation signature at {start}")
                b, q, end = b, 4, length - 4
            else:
                b, q, end = self.read(start, length), 0, length
            hsize = 4 + (2 if track_order else 0)
            while q + hsize <= end:
                mtype, msize, mflags = b[q], struct.unpack_from("<H", b, q + 1)[0], b[q + 3]
                body = b[q + hsize:q + hsize + msize]
                if mtype == _CONTINUATION:
                    chunks.append((self._uint(body, 0, self.so), self._uint(body, self.so, self.sl), True))
                out.append((mtype, mflags, body))
                q += hsize + msize
        return out

    # --- datatypes and dataspaces -------------------------------------------------------

    def datatype(self, b: bytes, p: int = 0):
        """(numpy dtype, bytes consumed, kind) of the datatype message at b[p:]; kind is one of
        'number', 'string', 'bool', 'complex', 'compound'."""
        cv = b[p]
        cls, version = cv & 0x0F, cv >> 4
        bits = b[p + 1] | (b[p + 2] << 8) | (b[p + 3] << 16)
        size = struct.unpack_from("<I", b, p + 4)[0]
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}"), 12, "number"
        if cls == 1:
            return np.dtype(f"{order}f{size}"), 20, "number"
        if cls == 3:
            return np.dtype(f"S{size}"), 8, "string"
        if cls == 4:
            return np.dtype(f"{order}u{size}"), 12, ("bool" if size == 1 else "number")
        if cls == 6:
            nmembers = bits & 0xFFFF
            q = p + 8
            names, formats, offsets = [], [], []
            for _ in range(nmembers):
                end = b.index(b"\0", q)
                name = b[q:end].decode("utf-8")
                if version < 3:
                    q += _pad8(end - q + 1)
                    offset = struct.unpack_from("<I", b, q)[0]
                    q += 4
                    if version == 1:
                        q += 28                            # dimensionality, permutation, sizes
                else:
                    q = end + 1
                    width = 1 if size < 256 else 2 if size < 65536 else 3 if size < 1 << 24 else 4
                    offset = int.from_bytes(b[q:q + width], "little")
                    q += width
                mdt, used, _ = self.datatype(b, q)
                q += used
                names.append(name)
                formats.append(mdt)
                offsets.append(offset)
            dt = np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})
            if (nmembers == 2 and names in (["r", "i"], ["real", "imag"]) and formats[0] == formats[1]
                    and formats[0].kind == "f" and offsets == [0, formats[0].itemsize]):
                return dt, q - p, "complex"
            return dt, q - p, "compound"
        if cls == 8:
            base, used, _ = self.datatype(b, p + 8)
            nmembers = bits & 0xFFFF
            q = p + 8 + used
            names = []
            for _ in range(nmembers):
                end = b.index(b"\0", q)
                names.append(b[q:end].decode("utf-8"))
                q = end + 1 if version >= 3 else q + _pad8(end - q + 1)
            q += nmembers * base.itemsize
            kind = "bool" if sorted(names) == ["FALSE", "TRUE"] else "number"
            return base, q - p, kind
        raise NotImplementedError(f"HDF5 reader: datatype class {cls} (variable-length, reference, opaque, "
                                  "array or time data) is not read here")

    def dataspace(self, b: bytes) -> tuple:
        version, rank, flags = b[0], b[1], b[2]
        if version == 1:
            p = 8
        elif version == 2:
            if b[3] == 2:                                   # the null dataspace
                return None
            p = 4
        else:
            raise FormatError(f"HDF5 reader: dataspace version {version} is not known")
        return tuple(self._uint(b, p + i * self.sl, self.sl) for i in range(rank))

    def _decode(self, raw: bytes, dt: np.dtype, kind: str, shape, python_scalar: bool = True):
        if shape is None:
            return None
        n = int(np.prod(shape)) if shape else 1
        a = np.frombuffer(raw[:n * dt.itemsize], dtype=dt, count=n)
        return self._convert(a, kind, shape, python_scalar)

    def _convert(self, a: np.ndarray, kind: str, shape, python_scalar: bool = True):
        """The decoded array of `shape`; a scalar becomes a Python value when `python_scalar`
        (attributes), and strings become str."""
        if kind == "complex":
            part = a.dtype.fields[a.dtype.names[0]][0]
            c = np.empty(a.shape, dtype=np.result_type(part, np.complex64))
            c.real, c.imag = a[a.dtype.names[0]], a[a.dtype.names[1]]
            a = c
        elif kind == "bool":
            a = a != 0
        elif a.dtype.kind in "iuf" and a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        a = a.reshape(shape)
        if kind == "string":
            def text(x):
                return bytes(x).split(b"\0", 1)[0].decode("utf-8")
            return text(a[()]) if a.ndim == 0 else np.array([text(x) for x in a.ravel()], dtype=object).reshape(shape)
        if a.ndim == 0 and python_scalar:
            return a[()].item()
        return np.array(a)

    def attribute(self, b: bytes):
        version = b[0]
        nsize, tsize, ssize = struct.unpack_from("<HHH", b, 2)
        if version == 1:
            p = 8
            name = b[p:p + nsize].split(b"\0", 1)[0].decode("utf-8")
            p += _pad8(nsize)
            dt, _, kind = self.datatype(b, p)
            p += _pad8(tsize)
            shape = self.dataspace(b[p:p + ssize])
            p += _pad8(ssize)
        elif version in (2, 3):
            p = 8 if version == 2 else 9
            name = b[p:p + nsize].split(b"\0", 1)[0].decode("utf-8")
            p += nsize
            if b[1] & 0x03:
                raise NotImplementedError("HDF5 reader: shared attribute datatypes are not read here")
            dt, _, kind = self.datatype(b, p)
            p += tsize
            shape = self.dataspace(b[p:p + ssize])
            p += ssize
        else:
            raise FormatError(f"HDF5 reader: attribute message version {version} is not known")
        return name, self._decode(b[p:], dt, kind, shape)

    # --- objects ------------------------------------------------------------------------

    def obj(self, addr: int):
        msgs = self.messages(addr)
        types = {t for t, _, _ in msgs}
        attrs = {}
        for t, flags, body in msgs:
            if t == _ATTRIBUTE:
                k, v = self.attribute(body)
                attrs[k] = v
            elif t == _ATTR_INFO:
                heap = self._uint(body, 2 + (2 if body[1] & 1 else 0), self.so)
                if not self._undef(heap, self.so):
                    raise NotImplementedError("HDF5 reader: dense attribute storage is not read here")
        if _LAYOUT in types:
            ds = Dataset(self._dataset(msgs), attrs)
            return ds
        g = Group()
        g.attrs = attrs
        for t, flags, body in msgs:
            if t == _SYMBOL_TABLE:
                btree, heap = self._uint(body, 0, self.so), self._uint(body, self.so, self.so)
                for name, child in self._symbol_table(btree, heap):
                    g.members[name] = self.obj(child)
            elif t == _LINK:
                name, child = self._link(body)
                if child is not None:
                    g.members[name] = self.obj(child)
            elif t == _LINK_INFO:
                heap = self._uint(body, 2 + (8 if body[1] & 1 else 0), self.so)
                if not self._undef(heap, self.so):
                    raise NotImplementedError("HDF5 reader: dense link storage (a group of many members "
                                              "written under the latest format) is not read here")
        return g

    def _dataset(self, msgs) -> np.ndarray:
        shape = dt = kind = layout = None
        for t, flags, body in msgs:
            if t == _DATASPACE:
                shape = self.dataspace(body)
            elif t == _DATATYPE:
                if flags & 0x02:
                    raise NotImplementedError("HDF5 reader: committed (shared) datatypes are not read here")
                dt, _, kind = self.datatype(body)
            elif t == _LAYOUT:
                layout = body
        if shape is None:
            return None
        n = int(np.prod(shape)) if shape else 1
        version = layout[0]
        if version in (1, 2):
            rank, cls = layout[1], layout[2]
            p = 8
            addr = self._uint(layout, p, self.so) if cls != 0 else None
            p += (self.so if cls != 0 else 0) + 4 * rank
            if cls == 0:
                size = struct.unpack_from("<I", layout, p)[0]
                raw = layout[p + 4:p + 4 + size]
        elif version in (3, 4):
            cls = layout[1]
            if cls == 0:
                size = struct.unpack_from("<H", layout, 2)[0]
                raw = layout[4:4 + size]
            elif cls == 1:
                addr = self._uint(layout, 2, self.so)
            else:
                cls = 2
        else:
            raise FormatError(f"HDF5 reader: data layout version {version} is not known")
        if cls == 2:
            raise NotImplementedError("HDF5 reader: chunked datasets are not read here (write contiguous data)")
        if cls == 1:
            if n == 0 or self._undef(addr, self.so):
                raw = b"\0" * (n * dt.itemsize)
            else:
                self.f.seek(self.base + addr)
                a = np.fromfile(self.f, dtype=dt, count=n)
                if a.size != n:
                    raise FormatError("HDF5 reader: the file ends inside a dataset")
                return self._convert(a, kind, shape, python_scalar=False)
        return self._decode(raw, dt, kind, shape, python_scalar=False)

    def _link(self, b: bytes):
        flags = b[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = b[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        width = 1 << (flags & 3)
        n = self._uint(b, p, width)
        p += width
        name = b[p:p + n].decode("utf-8")
        p += n
        if ltype != 0:
            return name, None                       # soft and external links are not followed
        return name, self._uint(b, p, self.so)

    def _symbol_table(self, btree: int, heap: int) -> list:
        h = self.read(heap, 32)
        if h[:4] != b"HEAP":
            raise FormatError(f"HDF5 reader: no local heap at {heap}")
        dsize = self._uint(h, 8, self.sl)
        daddr = self._uint(h, 8 + 2 * self.sl, self.so)
        names = self.read(daddr, dsize)
        out = []
        for snod in self._btree_leaves(btree):
            s = self.read(snod, 8)
            if s[:4] != b"SNOD":
                raise FormatError(f"HDF5 reader: no symbol-table node at {snod}")
            count = struct.unpack_from("<H", s, 6)[0]
            esize = self.so + self.sl + 24
            body = self.read(snod + 8, count * esize)
            for i in range(count):
                e = body[i * esize:(i + 1) * esize]
                off = self._uint(e, 0, self.sl)
                name = names[off:names.index(b"\0", off)].decode("utf-8")
                out.append((name, self._uint(e, self.sl, self.so)))
        return out

    def _btree_leaves(self, addr: int) -> list:
        h = self.read(addr, 8 + 2 * self.so)
        if h[:4] != b"TREE" or h[4] != 0:
            raise FormatError(f"HDF5 reader: no group B-tree node at {addr}")
        level, used = h[5], struct.unpack_from("<H", h, 6)[0]
        body = self.read(addr + 8 + 2 * self.so, used * (self.sl + self.so) + self.sl)
        children = [self._uint(body, self.sl + i * (self.sl + self.so), self.so) for i in range(used)]
        if level == 0:
            return children
        return [leaf for c in children for leaf in self._btree_leaves(c)]


def read_file(path) -> Group:
    """The root group of the HDF5 file at `path`, every dataset read into memory."""
    with open(path, "rb") as f:
        r = _Reader(f)
        return r.obj(r.root_header)


def is_hdf5(path) -> bool:
    """True when the file at `path` carries the HDF5 signature at offset 0 or after a user block."""
    with open(path, "rb") as f:
        at = 0
        while True:
            f.seek(at)
            head = f.read(8)
            if len(head) < 8:
                return False
            if head == SIGNATURE:
                return True
            at = 512 if at == 0 else 2 * at
