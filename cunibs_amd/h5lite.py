"""A read-only subset of HDF5, enough for coil files and field-result files without h5py.

The reference stores coils as one ``dipoles`` dataset with string/number attributes
(src/cunibs/coil.py:66-98, written by ``encode_ccd`` with gzip compression) and opens them with
h5py. This module walks the structures the public HDF5 File Format Specification (version 3.0)
defines for files that h5py writes with its default settings, plus the newer object-header forms:

* superblock versions 0-3; object headers version 1 and 2 (with continuation blocks);
* groups as symbol tables (v1 B-tree + local heap) or as compact link messages;
* datasets with compact, contiguous or chunked (v1 B-tree) layout, filters deflate, shuffle and
  Fletcher-32 (checksum stripped, not verified);
* datatypes fixed-point, floating-point, fixed strings and variable-length strings (global heap);
* attributes of those types, scalar or simple dataspaces.

Anything else raises ``NotImplementedError`` naming the structure, never a silent misread.

``write_file`` writes the matching subset: superblock version 0, version-1 object headers,
symbol-table groups, contiguous datasets of little-endian integers and IEEE floats, and
attributes of those types or of variable-length UTF-8 strings (one global heap collection). Each
structure is encoded as the HDF5 library encodes it in files h5py writes with its defaults (the
datatype, dataspace, fill-value and attribute encodings follow such a file byte for byte).
"""
from __future__ import annotations

import zlib
from pathlib import Path
from typing import NamedTuple

import numpy as np

_SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEFINED = 0xFFFFFFFFFFFFFFFF


class _VlenString:
    """Marker type of a variable-length string datatype."""


class _Reader:
    def __init__(self, data: bytes, name: str) -> None:
        self.d = data
        self.name = name
        self._heaps: dict[int, dict[int, bytes]] = {}
        for start in (0, 512, 1024, 2048, 4096):
            if data[start:start + 8] == _SIGNATURE:
                break
        else:
            raise ValueError(f"{name}: not an HDF5 file (no superblock signature)")
        ver = data[start + 8]
        if ver in (0, 1):
            self.so, self.sl = data[start + 13], data[start + 14]
            p = start + 24 + (4 if ver == 1 else 0)
            self.base = self.addr(p)
            p += 4 * self.so  # base, free-space info, end of file, driver info
            self.root = self.addr(p + self.so)  # root symbol table entry: name offset, header address
        elif ver in (2, 3):
            self.so, self.sl = data[start + 9], data[start + 10]
            p = start + 12
            self.base = self.addr(p)
            self.root = self.addr(p + 3 * self.so)  # base, extension, end of file, root header
        else:
            raise NotImplementedError(f"{name}: HDF5 superblock version {ver}")
        if self.base == _UNDEFINED:
            self.base = 0

    # -- primitive fields -------------------------------------------------------------------
    def uint(self, p: int, width: int) -> int:
        return int.from_bytes(self.d[p:p + width], "little")

    def addr(self, p: int) -> int:
        return self.uint(p, self.so)

    def length(self, p: int) -> int:
        return self.uint(p, self.sl)

    def at(self, a: int) -> int:
        return self.base + a

    # -- object headers -----------------------------------------------------------------------
    def messages(self, a: int) -> list[tuple[int, bytes]]:
        p = self.at(a)
        out: list[tuple[int, bytes]] = []
        if self.d[p:p + 4] == b"OHDR":
            ver, flags = self.d[p + 4], self.d[p + 5]
            if ver != 2:
                raise NotImplementedError(f"{self.name}: object header version {ver}")
            q = p + 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
            width = 1 << (flags & 3)
            size = self.uint(q, width)
            q += width
            blocks = [(q, q + size)]
            creation = bool(flags & 0x04)
            while blocks:
                s, e = blocks.pop(0)
                hdr = 4 + (2 if creation else 0)
                while s + hdr <= e:
                    mtype, msize = self.d[s], self.uint(s + 1, 2)
                    body = self.d[s + hdr:s + hdr + msize]
                    s += hdr + msize
                    if mtype == 0x10:
                        ca = int.from_bytes(body[:self.so], "little")
                        cl = int.from_bytes(body[self.so:self.so + self.sl], "little")
                        c = self.at(ca)
                        if self.d[c:c + 4] != b"OCHK":
                            raise ValueError(f"{self.name}: bad object header continuation at {ca}")
                        blocks.append((c + 4, c + cl - 4))
                    elif mtype:
                        out.append((mtype, body))
            return out
        ver = self.d[p]
        if ver != 1:
            raise NotImplementedError(f"{self.name}: object header version {ver} at {a}")
        n_msgs = self.uint(p + 2, 2)
        blocks = [(p + 16, p + 16 + self.uint(p + 8, 4))]
        while blocks and len(out) < n_msgs:
            s, e = blocks.pop(0)
            while s + 8 <= e:
                mtype, msize = self.uint(s, 2), self.uint(s + 2, 2)
                body = self.d[s + 8:s + 8 + msize]
                s += 8 + msize
                if mtype == 0x10:
                    ca = int.from_bytes(body[:self.so], "little")
                    cl = int.from_bytes(body[self.so:self.so + self.sl], "little")
                    blocks.append((self.at(ca), self.at(ca) + cl))
                elif mtype:
                    out.append((mtype, body))
        return out

    # -- groups -------------------------------------------------------------------------------
    def links(self, msgs: list[tuple[int, bytes]]) -> dict[str, int]:
        names: dict[str, int] = {}
        for mtype, body in msgs:
            if mtype == 0x11:  # symbol table: v1 B-tree of symbol nodes, local heap of names
                btree = int.from_bytes(body[:self.so], "little")
                heap = int.from_bytes(body[self.so:2 * self.so], "little")
                self._symbol_tree(btree, self._local_heap(heap), names)
            elif mtype == 0x06:
                name, target = self._link(body)
                if target is not None:
                    names[name] = target
            elif mtype == 0x02:
                fheap = int.from_bytes(body[2 + (8 if body[1] & 1 else 0):][:self.so], "little")
                if fheap != _UNDEFINED:
                    raise NotImplementedError(f"{self.name}: dense link storage (fractal heap)")
        return names

    def _local_heap(self, a: int) -> int:
        p = self.at(a)
        if self.d[p:p + 4] != b"HEAP":
            raise ValueError(f"{self.name}: bad local heap at {a}")
        return self.at(self.addr(p + 8 + 2 * self.sl))

    def _cstring(self, p: int) -> str:
        end = self.d.index(b"\0", p)
        return self.d[p:end].decode()

    def _symbol_tree(self, a: int, heap: int, names: dict[str, int]) -> None:
        p = self.at(a)
        if self.d[p:p + 4] != b"TREE" or self.d[p + 4] != 0:
            raise ValueError(f"{self.name}: bad group B-tree at {a}")
        level, used = self.d[p + 5], self.uint(p + 6, 2)
        q = p + 8 + 2 * self.so + self.sl  # past the siblings and the first key
        for _ in range(used):
            child = self.addr(q)
            q += self.so + self.sl
            if level:
                self._symbol_tree(child, heap, names)
                continue
            s = self.at(child)
            if self.d[s:s + 4] != b"SNOD":
                raise ValueError(f"{self.name}: bad symbol node at {child}")
            entry = 2 * self.so + 24
            for i in range(self.uint(s + 6, 2)):
                e = s + 8 + i * entry
                names[self._cstring(heap + self.addr(e))] = self.addr(e + self.so)

    def _link(self, body: bytes) -> tuple[str, int | None]:
        flags = body[1]
        q = 2
        ltype = 0
        if flags & 0x08:
            ltype = body[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        width = 1 << (flags & 3)
        n = int.from_bytes(body[q:q + width], "little")
        q += width
        name = body[q:q + n].decode()
        q += n
        if ltype != 0:
            return name, None  # soft and external links are not followed
        return name, int.from_bytes(body[q:q + self.so], "little")

    # -- datatypes, dataspaces, attributes ----------------------------------------------------
    def datatype(self, b: bytes):
        cls, bits, size = b[0] & 0x0F, b[1] | b[2] << 8 | b[3] << 16, int.from_bytes(b[4:8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}")
        if cls == 1:
            if bits & 0x40:
                raise NotImplementedError(f"{self.name}: VAX floating-point order")
            return np.dtype(f"{order}f{size}")
        if cls == 3:
            return np.dtype(f"S{size}")
        if cls == 9 and (bits & 0x0F) == 1:
            return _VlenString
        raise NotImplementedError(f"{self.name}: HDF5 datatype class {cls}")

    def dataspace(self, b: bytes) -> tuple[int, ...] | None:
        ver, rank, flags = b[0], b[1], b[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            if b[3] == 2:
                return None  # null dataspace
            p = 4
        else:
            raise NotImplementedError(f"{self.name}: dataspace version {ver}")
        del flags
        return tuple(int.from_bytes(b[p + i * self.sl:p + (i + 1) * self.sl], "little") for i in range(rank))

    def _global_object(self, collection: int, index: int) -> bytes:
        if collection not in self._heaps:
            p = self.at(collection)
            if self.d[p:p + 4] != b"GCOL":
                raise ValueError(f"{self.name}: bad global heap at {collection}")
            end = p + self.length(p + 8)
            q = p + 8 + self.sl
            objs: dict[int, bytes] = {}
            while q + 8 + self.sl <= end:
                idx = self.uint(q, 2)
                if idx == 0:
                    break
                size = self.length(q + 8)
                objs[idx] = self.d[q + 8 + self.sl:q + 8 + self.sl + size]
                q += 8 + self.sl + (size + 7) // 8 * 8
            self._heaps[collection] = objs
        return self._heaps[collection][index]

    def values(self, dtype, shape: tuple[int, ...] | None, raw: bytes):
        if shape is None:
            return None
        count = int(np.prod(shape)) if shape else 1
        if dtype is _VlenString:
            step = 4 + self.so + 4
            items = []
            for i in range(count):
                e = raw[i * step:(i + 1) * step]
                n = int.from_bytes(e[:4], "little")
                col = int.from_bytes(e[4:4 + self.so], "little")
                idx = int.from_bytes(e[4 + self.so:], "little")
                items.append(self._global_object(col, idx)[:n].decode() if n else "")
            return items[0] if not shape else np.array(items, dtype=object).reshape(shape)
        arr = np.frombuffer(raw[:count * dtype.itemsize], dtype=dtype).reshape(shape)
        if dtype.kind == "S":
            arr = np.char.rstrip(arr, b"\0")  # null-terminated and null-padded forms alike
        arr = arr.astype(dtype.newbyteorder("="), copy=True) if dtype.kind in "iuf" else arr.copy()
        return arr[()] if not shape else arr

    def attributes(self, msgs: list[tuple[int, bytes]]) -> dict[str, object]:
        attrs: dict[str, object] = {}
        for mtype, b in msgs:
            if mtype != 0x0C:
                continue
            ver = b[0]
            n_name, n_type, n_space = (int.from_bytes(b[i:i + 2], "little") for i in (2, 4, 6))

            def pad(n: int) -> int:
                return (n + 7) // 8 * 8 if ver == 1 else n

            p = 8 if ver in (1, 2) else 9
            if ver not in (1, 2, 3):
                raise NotImplementedError(f"{self.name}: attribute message version {ver}")
            name = b[p:p + n_name].split(b"\0", 1)[0].decode()
            p += pad(n_name)
            dtype = self.datatype(b[p:p + n_type])
            p += pad(n_type)
            shape = self.dataspace(b[p:p + n_space])
            p += pad(n_space)
            attrs[name] = self.values(dtype, shape, b[p:])
        return attrs

    # -- dataset storage ----------------------------------------------------------------------
    def read_dataset(self, msgs: list[tuple[int, bytes]]) -> np.ndarray:
        found = {t: b for t, b in msgs}
        shape = self.dataspace(found[0x01])
        dtype = self.datatype(found[0x03])
        if dtype is _VlenString:
            raise NotImplementedError(f"{self.name}: variable-length string datasets")
        layout = found[0x08]
        if layout[0] != 3:
            raise NotImplementedError(f"{self.name}: data layout message version {layout[0]}")
        filters = self._filters(found.get(0x0B))
        if shape is None:
            return np.empty(0, dtype=dtype)
        count = int(np.prod(shape)) if shape else 1
        cls = layout[1]
        if cls == 0:
            n = int.from_bytes(layout[2:4], "little")
            raw = layout[4:4 + n]
        elif cls == 1:
            a = int.from_bytes(layout[2:2 + self.so], "little")
            raw = b"" if a == _UNDEFINED else self.d[self.at(a):self.at(a) + count * dtype.itemsize]
            if len(raw) < count * dtype.itemsize:
                raw = raw + bytes(count * dtype.itemsize - len(raw))  # never written: fill value 0
        elif cls == 2:
            return self._chunked(layout, shape, dtype, filters)
        else:
            raise NotImplementedError(f"{self.name}: data layout class {cls}")
        return self.values(dtype, shape, raw) if shape else np.asarray(self.values(dtype, shape, raw))

    def _filters(self, b: bytes | None) -> list[tuple[int, list[int]]]:
        if b is None:
            return []
        ver, n = b[0], b[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(n):
            fid = int.from_bytes(b[p:p + 2], "little")
            name_len = int.from_bytes(b[p + 2:p + 4], "little") if (ver == 1 or fid >= 256) else 0
            p += 2 + (2 if (ver == 1 or fid >= 256) else 0)
            n_vals = int.from_bytes(b[p + 2:p + 4], "little")
            p += 4
            p += (name_len + 7) // 8 * 8 if ver == 1 else name_len
            vals = [int.from_bytes(b[p + 4 * i:p + 4 * i + 4], "little") for i in range(n_vals)]
            p += 4 * n_vals + (4 if ver == 1 and n_vals % 2 else 0)
            out.append((fid, vals))
        return out

    def _unfilter(self, raw: bytes, filters, mask: int, itemsize: int) -> bytes:
        for i in range(len(filters) - 1, -1, -1):
            if mask & (1 << i):
                continue
            fid, _ = filters[i]
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                a = np.frombuffer(raw, dtype=np.uint8)
                raw = a.reshape(itemsize, -1).T.tobytes()
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise NotImplementedError(f"{self.name}: HDF5 filter {fid}")
        return raw

    def _chunked(self, layout: bytes, shape, dtype, filters) -> np.ndarray:
        rank = layout[2] - 1
        btree = int.from_bytes(layout[3:3 + self.so], "little")
        q = 3 + self.so
        chunk = tuple(int.from_bytes(layout[q + 4 * i:q + 4 * i + 4], "little") for i in range(rank))
        out = np.zeros(shape, dtype=dtype.newbyteorder("="))
        if btree == _UNDEFINED:
            return out
        stack = [btree]
        key = 8 + 8 * (rank + 1)
        while stack:
            p = self.at(stack.pop())
            if self.d[p:p + 4] != b"TREE" or self.d[p + 4] != 1:
                raise ValueError(f"{self.name}: bad chunk B-tree")
            level, used = self.d[p + 5], self.uint(p + 6, 2)
            q = p + 8 + 2 * self.so
            for _ in range(used):
                size, mask = self.uint(q, 4), self.uint(q + 4, 4)
                offs = [self.uint(q + 8 + 8 * i, 8) for i in range(rank)]
                child = self.addr(q + key)
                q += key + self.so
                if level:
                    stack.append(child)
                    continue
                raw = self._unfilter(self.d[self.at(child):self.at(child) + size], filters, mask, dtype.itemsize)
                block = np.frombuffer(raw[:int(np.prod(chunk)) * dtype.itemsize], dtype=dtype).reshape(chunk)
                dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, chunk, shape))
                src = tuple(slice(0, d.stop - d.start) for d in dst)
                out[dst] = block[src]
        return out


class Node:
    """A group or dataset of an opened file."""

    def __init__(self, reader: _Reader, addr: int, path: str) -> None:
        self._r = reader
        self._msgs = reader.messages(addr)
        self.name = path
        self.attrs = reader.attributes(self._msgs)
        types = {t for t, _ in self._msgs}
        self.is_dataset = 0x08 in types
        self._links = None if self.is_dataset else reader.links(self._msgs)

    def keys(self) -> list[str]:
        if self._links is None:
            raise TypeError(f"{self.name} is a dataset")
        return list(self._links)

    def __contains__(self, name: str) -> bool:
        try:
            self[name]
        except KeyError:
            return False
        return True

    def __getitem__(self, name: str) -> Node:
        node = self
        for part in [p for p in name.split("/") if p]:
            if node._links is None or part not in node._links:
                raise KeyError(f"{name!r} not found in {self.name}")
            node = Node(self._r, node._links[part], f"{node.name.rstrip('/')}/{part}")
        return node

    def read(self) -> np.ndarray:
        if not self.is_dataset:
            raise TypeError(f"{self.name} is a group")
        return self._r.read_dataset(self._msgs)

    def __array__(self, dtype=None, copy=None):
        a = self.read()
        return a if dtype is None else a.astype(dtype)

    @property
    def shape(self) -> tuple[int, ...]:
        """From the dataspace message alone; the data is not read."""
        if not self.is_dataset:
            raise TypeError(f"{self.name} is a group")
        space = self._r.dataspace(next(b for t, b in self._msgs if t == 0x01))
        return (0,) if space is None else space


def open_file(path: str | Path) -> Node:
    """Open an HDF5 file for reading; the result is its root group."""
    p = Path(path)
    r = _Reader(p.read_bytes(), str(p))
    return Node(r, r.root, "/")


# ------------------------------------------------------------------------------------------------
# Writer
# ------------------------------------------------------------------------------------------------

_ATTRS = "__attrs__"
_FREE_NULL = 1  # a local heap's "no free block" offset, as the HDF5 library stores it
_INTERNAL_K = 16
_VLEN_UTF8 = bytes.fromhex("1901010010000000100000000100000000000800")
_DEFLATE_PIPELINE = bytes.fromhex("010100000000000001000800010001006465666c617465000400000000000000")
_CHUNK_K = 32  # the B-tree K of chunk indexes (a version-0 superblock does not store it)


class DatasetSpec(NamedTuple):
    """A dataset to write together with its attributes."""

    data: np.ndarray
    attrs: dict


def _pad8(b: bytes) -> bytes:
    return b + bytes(-len(b) % 8)


def _u(v: int, width: int) -> bytes:
    return int(v).to_bytes(width, "little")


def _numeric_type(dt: np.dtype) -> bytes:
    dt = np.dtype(dt)
    if dt.kind in "iu":
        return bytes([0x10, 0x08 if dt.kind == "i" else 0x00, 0, 0]) + _u(dt.itemsize, 4) + _u(0, 2) + \
            _u(8 * dt.itemsize, 2)
    if dt == np.float64:
        return bytes.fromhex("11203f00") + _u(8, 4) + _u(0, 2) + _u(64, 2) + bytes([52, 11, 0, 52]) + _u(1023, 4)
    if dt == np.float32:
        return bytes.fromhex("11201f00") + _u(4, 4) + _u(0, 2) + _u(32, 2) + bytes([23, 8, 0, 23]) + _u(127, 4)
    raise NotImplementedError(f"writing dtype {dt}")


def _dataspace(shape: tuple[int, ...]) -> bytes:
    return bytes([1, len(shape), 0, 0]) + bytes(4) + b"".join(_u(d, 8) for d in shape)


def _as_array(value) -> np.ndarray:
    a = np.asarray(value)
    if a.dtype == bool:
        a = a.astype(np.uint8)
    if a.dtype.kind in "iuf":
        return a.astype(a.dtype.newbyteorder("<"), order="C")  # keeps 0-d values 0-d
    raise NotImplementedError(f"writing values of dtype {a.dtype}")


class _Writer:
    def __init__(self, tree: dict, compress: bool) -> None:
        self.buf = bytearray(96)  # the superblock, filled in last
        self.compress = compress
        self.strings: dict[str, int] = {}
        self._collect(tree)
        self.leaf_k = max(4, (self._widest(tree) + 1) // 2)
        self.gcol = self._global_heap() if self.strings else 0
        root, btree, heap = self._group(tree)
        eof = len(self.buf)
        sb = _SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + _u(self.leaf_k, 2) + _u(_INTERNAL_K, 2) + _u(0, 4)
        sb += _u(0, 8) + _u(_UNDEFINED, 8) + _u(eof, 8) + _u(_UNDEFINED, 8)
        sb += _u(0, 8) + _u(root, 8) + _u(1, 4) + _u(0, 4) + _u(btree, 8) + _u(heap, 8)
        assert len(sb) == 96
        self.buf[:96] = sb

    def _collect(self, node) -> None:
        if isinstance(node, (dict, DatasetSpec)):
            attrs = node.attrs if isinstance(node, DatasetSpec) else node.get(_ATTRS, {})
            for k, v in attrs.items():
                for item in ([v] if isinstance(v, str) else (v if isinstance(v, (list, tuple)) and v
                                                              and all(isinstance(x, str) for x in v) else [])):
                    self.strings.setdefault(item, len(self.strings) + 1)
        if isinstance(node, dict):
            for k, v in node.items():
                if k != _ATTRS:
                    self._collect(v)

    def _widest(self, node) -> int:
        if not isinstance(node, dict):
            return 0
        kids = [v for k, v in node.items() if k != _ATTRS]
        return max([len(kids)] + [self._widest(v) for v in kids])

    def put(self, data: bytes) -> int:
        addr = len(self.buf)
        self.buf += _pad8(bytes(data))
        return addr

    def _global_heap(self) -> int:
        objs = b""
        for text, idx in self.strings.items():
            raw = text.encode()
            objs += _u(idx, 2) + _u(0, 2) + bytes(4) + _u(len(raw), 8) + _pad8(raw)
        size = max(4096, 16 + len(objs) + 16)
        free = size - 16 - len(objs)
        body = b"GCOL" + bytes([1, 0, 0, 0]) + _u(size, 8) + objs + _u(0, 2) + _u(0, 2) + bytes(4) + _u(free, 8)
        return self.put(body + bytes(size - len(body)))

    def _attribute(self, name: str, value) -> bytes:
        if isinstance(value, str) or (isinstance(value, (list, tuple)) and value
                                      and all(isinstance(x, str) for x in value)):
            items = [value] if isinstance(value, str) else list(value)
            dtype, shape = _VLEN_UTF8, (() if isinstance(value, str) else (len(items),))
            data = b"".join(_u(len(t.encode()), 4) + _u(self.gcol, 8) + _u(self.strings[t], 4) for t in items)
        else:
            a = _as_array(value)
            dtype, shape, data = _numeric_type(a.dtype), a.shape, a.tobytes()
        nm = name.encode() + b"\0"
        space = _dataspace(shape)
        return bytes([1, 0]) + _u(len(nm), 2) + _u(len(dtype), 2) + _u(len(space), 2) + _pad8(nm) + \
            _pad8(dtype) + _pad8(space) + data

    def _header(self, messages: list[tuple[int, int, bytes]]) -> int:
        body = b""
        for mtype, flags, data in messages:
            data = _pad8(data)
            body += _u(mtype, 2) + _u(len(data), 2) + bytes([flags, 0, 0, 0]) + data
        return self.put(bytes([1, 0]) + _u(len(messages), 2) + _u(1, 4) + _u(len(body), 4) + bytes(4) + body)

    def _attr_messages(self, attrs: dict) -> list[tuple[int, int, bytes]]:
        return [(0x0C, 0, self._attribute(k, v)) for k, v in attrs.items()]

    def _dataset(self, value, attrs: dict | None = None) -> int:
        a = _as_array(value)
        raw = a.tobytes()
        msgs = [(0x01, 0, _dataspace(a.shape)), (0x03, 1, _numeric_type(a.dtype))]
        if self.compress and a.ndim >= 1 and a.size > 0:
            # One chunk spanning the array, deflated, indexed by a one-entry v1 B-tree.
            packed = zlib.compress(raw, 4)
            data = self.put(packed)
            key = lambda offs: _u(len(packed), 4) + _u(0, 4) + b"".join(_u(o, 8) for o in offs)
            node = b"TREE" + bytes([1, 0]) + _u(1, 2) + _u(_UNDEFINED, 8) + _u(_UNDEFINED, 8)
            node += key([0] * (a.ndim + 1)) + _u(data, 8) + key(list(a.shape) + [0])
            full = 24 + (2 * _CHUNK_K + 1) * (8 + 8 * (a.ndim + 1)) + 2 * _CHUNK_K * 8
            btree = self.put(node + bytes(full - len(node)))
            layout = bytes([3, 2, a.ndim + 1]) + _u(btree, 8) + b"".join(_u(d, 4) for d in a.shape) + \
                _u(a.dtype.itemsize, 4)
            msgs += [(0x05, 1, bytes([2, 3, 2, 1]) + _u(0, 4)), (0x0B, 1, _DEFLATE_PIPELINE), (0x08, 0, layout)]
        else:
            addr = self.put(raw) if raw else _UNDEFINED
            msgs += [(0x05, 1, bytes([2, 2, 2, 1]) + _u(0, 4)),
                     (0x08, 0, bytes([3, 1]) + _u(addr, 8) + _u(len(raw), 8))]
        return self._header(msgs + self._attr_messages(attrs or {}))

    def _group(self, node: dict) -> tuple[int, int, int]:
        names = sorted((k for k in node if k != _ATTRS), key=lambda k: k.encode())
        children = {}
        for k in names:
            v = node[k]
            if isinstance(v, dict):
                children[k] = self._group(v)[0]
            elif isinstance(v, DatasetSpec):
                children[k] = self._dataset(v.data, v.attrs)
            else:
                children[k] = self._dataset(v)
        heap_data, offsets = bytearray(8), {}
        for k in names:
            offsets[k] = len(heap_data)
            heap_data += _pad8(k.encode() + b"\0")
        seg = self.put(bytes(heap_data))
        heap = self.put(b"HEAP" + bytes(4) + _u(len(heap_data), 8) + _u(_FREE_NULL, 8) + _u(seg, 8))
        entry = lambda k: _u(offsets[k], 8) + _u(children[k], 8) + _u(0, 4) + _u(0, 4) + bytes(16)
        tree = b"TREE" + bytes([0, 0]) + _u(1 if names else 0, 2) + _u(_UNDEFINED, 8) + _u(_UNDEFINED, 8)
        if names:  # an empty group's B-tree has no entries and no symbol node
            snod = b"SNOD" + bytes([1, 0]) + _u(len(names), 2) + b"".join(entry(k) for k in names)
            snod_addr = self.put(snod + bytes(8 + 2 * self.leaf_k * 40 - len(snod)))
            tree += _u(0, 8) + _u(snod_addr, 8) + _u(offsets[names[-1]], 8)
        full = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8
        btree = self.put(tree + bytes(full - len(tree)))
        msgs = [(0x11, 0, _u(btree, 8) + _u(heap, 8))] + self._attr_messages(node.get(_ATTRS, {}))
        return self._header(msgs), btree, heap


def write_file(path: str | Path, tree: dict, compression: str | None = None) -> None:
    """Write an HDF5 file. ``tree`` maps member names to arrays (datasets) or to nested dicts
    (groups) or to ``DatasetSpec(array, attrs)``; the key ``"__attrs__"`` of a dict holds that
    group's attributes (numbers, arrays of numbers, strings or lists of strings).
    ``compression="gzip"`` stores each non-empty array as one deflated chunk."""
    if compression not in (None, "gzip"):
        raise ValueError(f"unsupported compression {compression!r}")
    Path(path).write_bytes(bytes(_Writer(tree, compression == "gzip").buf))
