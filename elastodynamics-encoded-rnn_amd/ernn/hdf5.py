"""Reader and writer for the HDF5 files of the reference's observed data,
without h5py.

The reference writes its FEM observed data with h5py
(Parameter_inversion/FEM_generate_data.py:902-945,
Damage_inversion/FEM_generate_data.py:1619-1684) and reads it back with h5py
(Parameter_inversion/Inversion_adam.py:536-550,
Damage_inversion/Inversion_adam.py:617-645): datasets ux, uy (nt_ds, nx, ny)
float32 in chunks of one frame (1, nx, ny), gzip level 4; x, y, t float32;
damage_mask uint8; groups source_info and damage_info (with region_<i>
subgroups) of scalars, small arrays and strings.  h5py with HDF5's default
format bounds, and libver='earliest' in the damage writer, produce the
version-0 layout of the HDF5 file format specification: a version-0
superblock, version-1 object headers, groups as symbol tables (a version-1
B-tree of symbol nodes plus a local heap), chunked data indexed by a
version-1 B-tree, variable-length strings in global heap collections.

`File` reads that layout (and the version-2 superblock, version-2 object
headers and compact link groups that newer format bounds write), with the
deflate, shuffle and Fletcher-32 filters, and exposes the part of h5py's
interface the reference's loaders use: `f[name][:]`, `f[name][()]`,
`name in f`, `group.keys()`, nested paths.  `write` produces version-0 files
of the same layout (what the reference's writers produce), readable by
h5py.  Structures this reader does not handle (dense link storage, the
chunk indexes of version-4 layouts other than single-chunk and implicit,
compound and reference datatypes, external storage, the szip filter)
raise `NotImplementedError` naming the structure.

Compression and decompression run on a thread pool (zlib releases the
GIL), one chunk per task; decompressed chunks go straight into the output
array.
"""
from __future__ import annotations

import os
import struct
import zlib
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Iterator, List, Optional, Sequence, Tuple, Union

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF

# message types (HDF5 file format specification, section IV.A.2)
MSG_NIL, MSG_DATASPACE, MSG_LINK_INFO, MSG_DATATYPE, MSG_FILL_OLD, MSG_FILL = 0x0, 0x1, 0x2, 0x3, 0x4, 0x5
MSG_LINK, MSG_EXTERNAL, MSG_LAYOUT, MSG_BOGUS, MSG_GROUP_INFO, MSG_FILTERS = 0x6, 0x7, 0x8, 0x9, 0xA, 0xB
MSG_ATTRIBUTE, MSG_COMMENT, MSG_MTIME_OLD, MSG_SHARED_TABLE, MSG_CONTINUATION = 0xC, 0xD, 0xE, 0xF, 0x10
MSG_SYMBOL_TABLE, MSG_MTIME = 0x11, 0x12

FILTER_DEFLATE, FILTER_SHUFFLE, FILTER_FLETCHER32, FILTER_SZIP = 1, 2, 3, 4


class HDF5Error(ValueError):
    """A file that is not HDF5 or is damaged."""


# ----------------------------------------------------------------------------- parsing helpers
class _Cursor:
    """Little-endian reads from a byte buffer at a moving position."""

    def __init__(self, buf, pos: int, so: int, sl: int):
        self.buf, self.pos, self.so, self.sl = buf, pos, so, sl

    def u(self, n: int) -> int:
        v = int.from_bytes(self.buf[self.pos:self.pos + n], "little")
        self.pos += n
        return v

    def off(self) -> int:
        return self.u(self.so)

    def length(self) -> int:
        return self.u(self.sl)

    def raw(self, n: int) -> bytes:
        v = bytes(self.buf[self.pos:self.pos + n])
        self.pos += n
        return v

    def skip(self, n: int) -> None:
        self.pos += n


def _align8(n: int) -> int:
    return (n + 7) & ~7


class _Datatype:
    """A parsed datatype message: numpy dtype, or a string kind."""

    def __init__(self, cls: int, size: int, dtype: Optional[np.dtype], kind: str = "numeric",
                 base: Optional["_Datatype"] = None):
        self.cls, self.size, self.dtype, self.kind, self.base = cls, size, dtype, kind, base


def _parse_datatype(c: _Cursor) -> _Datatype:
    b0 = c.u(1)
    cls, ver = b0 & 0x0F, b0 >> 4
    bits = c.u(3)
    size = c.u(4)
    if cls == 0:   # fixed-point
        c.skip(4)   # bit offset, precision
        if bits & 1:
            raise NotImplementedError("big-endian integer datatype")
        signed = bool(bits & 0x8)
        return _Datatype(cls, size, np.dtype(f"<{'i' if signed else 'u'}{size}"))
    if cls == 1:   # floating point
        c.skip(12)  # bit offset, precision, exponent / mantissa layout, bias
        if bits & 1 or bits & 0x40:
            raise NotImplementedError("non little-endian floating-point datatype")
        if size not in (2, 4, 8):
            raise NotImplementedError(f"{size}-byte floating-point datatype")
        return _Datatype(cls, size, np.dtype(f"<f{size}"))
    if cls == 3:   # fixed-length string
        return _Datatype(cls, size, np.dtype(f"S{size}"), kind="fstring")
    if cls == 8:   # enumeration (h5py's bool): values of the base integer type
        base = _parse_datatype(c)
        return _Datatype(cls, size, base.dtype, base=base)
    if cls == 9:   # variable length
        base = _parse_datatype(c)
        if bits & 0xF == 1:
            return _Datatype(cls, size, None, kind="vstring", base=base)
        return _Datatype(cls, size, None, kind="vlen", base=base)
    raise NotImplementedError(f"datatype class {cls}")


class _Node:
    """One object: its header messages, parsed on demand."""

    def __init__(self, f: "File", addr: int):
        self.f, self.addr = f, addr
        self.msgs: List[Tuple[int, int, int]] = []   # (type, data offset, size)
        f._read_header(addr, self.msgs)

    def find(self, mtype: int) -> List[Tuple[int, int]]:
        return [(o, n) for (t, o, n) in self.msgs if t == mtype]

    def first(self, mtype: int) -> Optional[Tuple[int, int]]:
        m = self.find(mtype)
        return m[0] if m else None


# ----------------------------------------------------------------------------- public objects
class Dataset:
    """h5py-style dataset: shape, dtype, [()] / [:] / [index] read."""

    def __init__(self, f: "File", node: _Node, name: str):
        self._f, self._node, self.name = f, node, name
        ds = node.first(MSG_DATASPACE)
        dt = node.first(MSG_DATATYPE)
        if ds is None or dt is None:
            raise HDF5Error(f"{name}: object has no dataspace or datatype message")
        self.shape = f._parse_dataspace(ds[0])
        self._type = _parse_datatype(_Cursor(f._buf, dt[0], f._so, f._sl))
        self.dtype = self._type.dtype if self._type.dtype is not None else np.dtype(object)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    def read(self) -> np.ndarray:
        return self._f._read_dataset(self)

    def __getitem__(self, key):
        a = self.read()
        if key == () or key is Ellipsis:
            return a[()] if a.shape == () else a
        return a[key]

    def __array__(self, dtype=None):
        a = self.read()
        return a if dtype is None else a.astype(dtype)

    def __len__(self) -> int:
        if not self.shape:
            raise TypeError("scalar dataset has no len()")
        return self.shape[0]

    def __repr__(self) -> str:
        return f"<HDF5 dataset {self.name!r}: shape {self.shape}, type {self.dtype}>"


class Group:
    """h5py-style group: keys(), [name], `name in group`, items()."""

    def __init__(self, f: "File", node: _Node, name: str):
        self._f, self._node, self.name = f, node, name
        self._links: Optional[Dict[str, int]] = None

    def _members(self) -> Dict[str, int]:
        if self._links is None:
            self._links = self._f._group_links(self._node, self.name)
        return self._links

    def keys(self) -> List[str]:
        return list(self._members().keys())

    def __iter__(self) -> Iterator[str]:
        return iter(self.keys())

    def __len__(self) -> int:
        return len(self._members())

    def __contains__(self, name: str) -> bool:
        try:
            self._resolve(name)
            return True
        except KeyError:
            return False

    def _resolve(self, path: str) -> int:
        node_addr = None
        g: Group = self
        parts = [p for p in path.split("/") if p]
        if not parts:
            return self._node.addr
        for i, p in enumerate(parts):
            links = g._members()
            if p not in links:
                raise KeyError(f"{path!r} not in {self.name!r}")
            node_addr = links[p]
            if i + 1 < len(parts):
                obj = self._f._object(node_addr, g.name.rstrip("/") + "/" + p)
                if not isinstance(obj, Group):
                    raise KeyError(f"{path!r}: {p!r} is not a group")
                g = obj
        return node_addr

    def __getitem__(self, path: str) -> Union["Group", Dataset]:
        addr = self._resolve(path)
        full = (self.name.rstrip("/") + "/" + path.strip("/")) if path.strip("/") else self.name
        return self._f._object(addr, full)

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def get(self, path: str, default=None):
        return self[path] if path in self else default

    def __repr__(self) -> str:
        return f"<HDF5 group {self.name!r} ({len(self)} members)>"


class File(Group):
    """An HDF5 file opened for reading (h5py.File(path, 'r') subset)."""

    def __init__(self, path: str, mode: str = "r", threads: Optional[int] = None):
        if mode != "r":
            raise ValueError("ernn.hdf5.File reads only; use ernn.hdf5.write to create files")
        self.filename = path
        with open(path, "rb") as fh:
            self._buf = memoryview(fh.read())
        self._threads = threads or min(16, os.cpu_count() or 1)
        self._nodes: Dict[int, _Node] = {}
        root = self._superblock()
        super().__init__(self, self._node_at(root), "/")

    # -- context manager, h5py style
    def __enter__(self) -> "File":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def close(self) -> None:
        self._buf = memoryview(b"")

    # -- superblock
    def _superblock(self) -> int:
        buf = self._buf
        base = None
        for cand in [0] + [512 << k for k in range(0, 16)]:
            if cand + 8 <= len(buf) and bytes(buf[cand:cand + 8]) == SIGNATURE:
                base = cand
                break
        if base is None:
            raise HDF5Error(f"{self.filename}: no HDF5 signature")
        ver = buf[base + 8]
        if ver in (0, 1):
            self._so, self._sl = buf[base + 13], buf[base + 14]
            c = _Cursor(buf, base + 24 + (4 if ver == 1 else 0), self._so, self._sl)
            self._base = c.off()
            c.off(); c.off(); c.off()   # free-space, end of file, driver block
            c.off()                    # root entry: link name offset
            root = c.off()             # root entry: object header address
        elif ver in (2, 3):
            self._so, self._sl = buf[base + 9], buf[base + 10]
            c = _Cursor(buf, base + 12, self._so, self._sl)
            self._base = c.off()
            c.off(); c.off()           # superblock extension, end of file
            root = c.off()
        else:
            raise NotImplementedError(f"superblock version {ver}")
        if self._base != 0 and self._base != base:
            raise NotImplementedError("superblock base address other than the signature's position")
        self._base = base if self._base == 0 else self._base
        return root

    def _cur(self, addr: int) -> _Cursor:
        return _Cursor(self._buf, self._base + addr, self._so, self._sl)

    # -- object headers
    def _node_at(self, addr: int) -> _Node:
        if addr not in self._nodes:
            self._nodes[addr] = _Node(self, addr)
        return self._nodes[addr]

    def _read_header(self, addr: int, msgs: List[Tuple[int, int, int]]) -> None:
        buf = self._buf
        p = self._base + addr
        if bytes(buf[p:p + 4]) == b"OHDR":
            self._read_header_v2(p, msgs)
            return
        if buf[p] != 1:
            raise HDF5Error(f"object header at {addr:#x}: version {buf[p]}")
        n_msgs = int.from_bytes(buf[p + 2:p + 4], "little")
        size = int.from_bytes(buf[p + 8:p + 12], "little")
        blocks = [(p + 16, size)]
        seen = 0
        while blocks and seen < n_msgs:
            start, length = blocks.pop(0)
            q, end = start, start + length
            while q + 8 <= end and seen < n_msgs:
                mtype = int.from_bytes(buf[q:q + 2], "little")
                msize = int.from_bytes(buf[q + 2:q + 4], "little")
                flags = buf[q + 4]
                data = q + 8
                seen += 1
                if flags & 0x2:
                    raise NotImplementedError("shared object header messages")
                if mtype == MSG_CONTINUATION:
                    c = _Cursor(buf, data, self._so, self._sl)
                    blocks.append((self._base + c.off(), c.length()))
                else:
                    msgs.append((mtype, data, msize))
                q = data + _align8(msize)

    def _read_header_v2(self, p: int, msgs: List[Tuple[int, int, int]]) -> None:
        buf = self._buf
        flags = buf[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16
        if flags & 0x10:
            q += 4
        nsz = 1 << (flags & 3)
        size0 = int.from_bytes(buf[q:q + nsz], "little")
        q += nsz
        blocks = [(q, size0)]
        creation_order = bool(flags & 0x4)
        while blocks:
            start, length = blocks.pop(0)
            r, end = start, start + length
            while r + 4 <= end:
                mtype = buf[r]
                msize = int.from_bytes(buf[r + 1:r + 3], "little")
                mflags = buf[r + 3]
                data = r + 4 + (2 if creation_order else 0)
                if data + msize > end:
                    break   # the gap before the checksum
                if mflags & 0x2:
                    raise NotImplementedError("shared object header messages")
                if mtype == MSG_CONTINUATION:
                    c = _Cursor(buf, data, self._so, self._sl)
                    a, n = c.off(), c.length()
                    a += self._base
                    if bytes(buf[a:a + 4]) != b"OCHK":
                        raise HDF5Error("object header continuation without OCHK signature")
                    blocks.append((a + 4, n - 8))
                elif mtype != MSG_NIL:
                    msgs.append((mtype, data, msize))
                r = data + msize

    def _object(self, addr: int, name: str) -> Union[Group, Dataset]:
        node = self._node_at(addr)
        if node.first(MSG_DATASPACE) is not None and node.first(MSG_LAYOUT) is not None:
            return Dataset(self, node, name)
        return Group(self, node, name)

    # -- groups
    def _group_links(self, node: _Node, name: str) -> Dict[str, int]:
        links: Dict[str, int] = {}
        st = node.first(MSG_SYMBOL_TABLE)
        if st is not None:
            c = _Cursor(self._buf, st[0], self._so, self._sl)
            btree, heap = c.off(), c.off()
            heap_data = self._local_heap(heap)
            for entry in self._btree_leaves(btree, group=True):
                self._symbol_node(entry, heap_data, links)
            return links
        for o, n in node.find(MSG_LINK):
            self._link_message(o, links)
        li = node.first(MSG_LINK_INFO)
        if li is not None:
            c = _Cursor(self._buf, li[0], self._so, self._sl)
            c.skip(1)
            fl = c.u(1)
            if fl & 1:
                c.skip(8)
            fheap = c.off()
            if fheap != UNDEF:
                raise NotImplementedError(f"{name}: dense link storage (fractal heap)")
        return links

    def _local_heap(self, addr: int) -> bytes:
        c = self._cur(addr)
        if c.raw(4) != b"HEAP":
            raise HDF5Error("local heap signature")
        c.skip(4)
        size = c.length()
        c.length()   # free list head
        data = c.off()
        p = self._base + data
        return bytes(self._buf[p:p + size])

    def _symbol_node(self, addr: int, heap: bytes, links: Dict[str, int]) -> None:
        c = self._cur(addr)
        if c.raw(4) != b"SNOD":
            raise HDF5Error("symbol table node signature")
        c.skip(2)
        n = c.u(2)
        for _ in range(n):
            name_off = c.off()
            obj = c.off()
            c.skip(8 + 16)   # cache type, reserved, scratch pad
            end = heap.index(b"\0", name_off)
            links[heap[name_off:end].decode("utf-8")] = obj

    def _link_message(self, off: int, links: Dict[str, int]) -> None:
        c = _Cursor(self._buf, off, self._so, self._sl)
        c.skip(1)
        flags = c.u(1)
        ltype = c.u(1) if flags & 0x8 else 0
        if flags & 0x4:
            c.skip(8)
        if flags & 0x10:
            c.skip(1)
        nlen = c.u(1 << (flags & 3))
        name = c.raw(nlen).decode("utf-8")
        if ltype != 0:
            raise NotImplementedError(f"link {name!r}: soft or external link")
        links[name] = c.off()

    # -- version-1 B-trees
    def _btree_leaves(self, addr: int, group: bool, ndims: int = 0) -> Iterator:
        """Group trees: child addresses (symbol nodes).  Chunk trees:
        (chunk address, stored size, filter mask, offsets)."""
        stack = [addr]
        while stack:
            a = stack.pop()
            c = self._cur(a)
            if c.raw(4) != b"TREE":
                raise HDF5Error("B-tree node signature")
            ntype, level, used = c.u(1), c.u(1), c.u(2)
            c.off(); c.off()   # siblings
            kids = []
            for _ in range(used):
                if ntype == 0:
                    c.length()
                    kids.append((c.off(), None))
                else:
                    size, mask = c.u(4), c.u(4)
                    offs = tuple(c.u(8) for _ in range(ndims + 1))
                    kids.append((c.off(), (size, mask, offs)))
            if level > 0:
                stack.extend(k for k, _ in reversed(kids))
            else:
                for child, key in kids:
                    yield child if group else (child, key[0], key[1], key[2][:ndims])

    # -- datasets
    def _parse_dataspace(self, off: int) -> Tuple[int, ...]:
        c = _Cursor(self._buf, off, self._so, self._sl)
        ver = c.u(1)
        rank = c.u(1)
        c.u(1)   # flags
        if ver == 1:
            c.skip(5)
        elif ver == 2:
            if c.u(1) == 2:
                return (0,)   # null dataspace: no elements
        else:
            raise NotImplementedError(f"dataspace version {ver}")
        return tuple(c.length() for _ in range(rank))

    def _filters(self, node: _Node) -> List[Tuple[int, int, Tuple[int, ...]]]:
        m = node.first(MSG_FILTERS)
        if m is None:
            return []
        c = _Cursor(self._buf, m[0], self._so, self._sl)
        ver, n = c.u(1), c.u(1)
        if ver == 1:
            c.skip(6)
        out = []
        for _ in range(n):
            fid = c.u(2)
            nlen = c.u(2) if (ver == 1 or fid >= 256) else 0
            flags, ncd = c.u(2), c.u(2)
            if nlen:
                c.skip(_align8(nlen) if ver == 1 else nlen)
            cd = tuple(c.u(4) for _ in range(ncd))
            if ver == 1 and ncd % 2:
                c.skip(4)
            out.append((fid, flags, cd))
        return out

    def _unfilter(self, raw: bytes, filters, mask: int) -> bytes:
        for k in range(len(filters) - 1, -1, -1):
            fid, flags, cd = filters[k]
            if mask & (1 << k):
                continue
            if fid == FILTER_DEFLATE:
                raw = zlib.decompress(raw)
            elif fid == FILTER_SHUFFLE:
                es = cd[0] if cd else 1
                a = np.frombuffer(raw, np.uint8)
                n = len(a) // es
                body = a[:n * es].reshape(es, n).T.reshape(-1)
                raw = body.tobytes() + bytes(a[n * es:])
            elif fid == FILTER_FLETCHER32:
                raw = raw[:-4]
            elif fid == FILTER_SZIP:
                raise NotImplementedError("szip filter")
            else:
                raise NotImplementedError(f"filter {fid}")
        return raw

    def _read_dataset(self, ds: Dataset) -> np.ndarray:
        node, t = ds._node, ds._type
        if node.first(MSG_EXTERNAL) is not None:
            raise NotImplementedError(f"{ds.name}: external storage")
        lay = node.first(MSG_LAYOUT)
        c = _Cursor(self._buf, lay[0], self._so, self._sl)
        ver = c.u(1)
        shape = ds.shape
        count = int(np.prod(shape, dtype=np.int64))
        esize = t.size
        if ver in (1, 2):
            rank = c.u(1)
            cls = c.u(1)
            c.skip(5)
            addr = c.off() if cls != 0 else None
            dims = tuple(c.u(4) for _ in range(rank))
            if cls == 0:
                n = c.u(4)
                return self._decode(t, bytes(self._buf[c.pos:c.pos + n]), shape)
            if cls == 1:
                return self._contiguous(t, addr, count * esize, shape)
            return self._chunked_v1(ds, addr, dims[:-1], shape)
        if ver == 3:
            cls = c.u(1)
            if cls == 0:
                n = c.u(2)
                return self._decode(t, bytes(self._buf[c.pos:c.pos + n]), shape)
            if cls == 1:
                addr = c.off()
                return self._contiguous(t, addr, c.length(), shape)
            if cls == 2:
                rank = c.u(1)
                addr = c.off()
                dims = tuple(c.u(4) for _ in range(rank))
                return self._chunked_v1(ds, addr, dims[:-1], shape)
            raise NotImplementedError(f"{ds.name}: layout class {cls}")
        if ver == 4:
            cls = c.u(1)
            if cls == 0:
                n = c.u(2)
                return self._decode(t, bytes(self._buf[c.pos:c.pos + n]), shape)
            if cls == 1:
                addr = c.off()
                return self._contiguous(t, addr, c.length(), shape)
            if cls == 2:
                flags = c.u(1)
                rank = c.u(1)
                enc = c.u(1)
                dims = tuple(c.u(enc) for _ in range(rank))
                itype = c.u(1)
                if itype == 1:   # single chunk
                    fsize, fmask = (c.length(), c.u(4)) if flags & 0x2 else (None, 0)
                    addr = c.off()
                    n = fsize if fsize is not None else count * esize
                    raw = self._unfilter(bytes(self._buf[self._base + addr:self._base + addr + n]),
                                         self._filters(node), fmask)
                    return self._decode(t, raw[:count * esize], shape)
                if itype == 2:   # implicit: unfiltered chunks stored in order
                    addr = c.off()
                    return self._implicit_chunks(ds, addr, dims[:-1], shape)
                if itype == 3:   # fixed array (fixed-size datasets under libver latest)
                    c.u(1)       # page bits (repeated in the array header)
                    addr = c.off()
                    return self._fixed_array_chunks(ds, addr, dims[:-1], shape)
                raise NotImplementedError(f"{ds.name}: version-4 chunk index type {itype} "
                                          "(extensible array or version-2 B-tree: datasets with unlimited dimensions)")
            raise NotImplementedError(f"{ds.name}: layout class {cls}")
        raise NotImplementedError(f"{ds.name}: layout version {ver}")

    def _contiguous(self, t: _Datatype, addr: int, nbytes: int, shape) -> np.ndarray:
        if addr == UNDEF:
            return self._fill(t, shape)
        p = self._base + addr
        return self._decode(t, bytes(self._buf[p:p + nbytes]), shape)

    def _fill(self, t: _Datatype, shape) -> np.ndarray:
        if t.dtype is None:
            return np.full(shape, b"" if t.kind == "vstring" else None, dtype=object)
        return np.zeros(shape, t.dtype)

    def _decode(self, t: _Datatype, raw: bytes, shape) -> np.ndarray:
        count = int(np.prod(shape, dtype=np.int64))
        if t.kind == "vstring" or t.kind == "vlen":
            out = np.empty(count, dtype=object)
            c = _Cursor(raw, 0, self._so, self._sl)
            for k in range(count):
                n = c.u(4)
                coll, idx = c.off(), c.u(4)
                data = self._global_heap(coll, idx) if coll not in (0, UNDEF) else b""
                if t.kind == "vstring":
                    out[k] = data[:n]
                else:
                    out[k] = np.frombuffer(data, t.base.dtype, count=n).copy()
            return out.reshape(shape)
        return np.frombuffer(raw, t.dtype, count=count).reshape(shape).copy()

    def _global_heap(self, coll: int, index: int) -> bytes:
        cache = self.__dict__.setdefault("_gcol", {})
        if coll not in cache:
            c = self._cur(coll)
            if c.raw(4) != b"GCOL":
                raise HDF5Error("global heap signature")
            c.skip(4)
            size = c.length()
            end = c.pos - 8 - self._sl + size
            objs = {}
            while c.pos + 8 + self._sl <= end:
                idx, _ref = c.u(2), c.u(2)
                c.skip(4)
                n = c.length()
                if idx == 0:
                    break
                objs[idx] = c.raw(n)
                c.skip(_align8(n) - n)
            cache[coll] = objs
        return cache[coll][index]

    def _chunk_grid(self, shape, cdims):
        """Copy box of one chunk (offsets) into the array: (output slices, chunk slices)."""
        def boxes(offs):
            out, src = [], []
            for o, n, s in zip(offs, cdims, shape):
                m = min(n, s - o)
                out.append(slice(o, o + m))
                src.append(slice(0, m))
            return tuple(out), tuple(src)
        return boxes

    def _chunked_v1(self, ds: Dataset, btree: int, cdims: Sequence[int], shape) -> np.ndarray:
        t = ds._type
        if t.dtype is None:
            raise NotImplementedError(f"{ds.name}: chunked variable-length data")
        out = np.zeros(shape, t.dtype)
        if btree == UNDEF:
            return out
        filters = self._filters(ds._node)
        rank = len(shape)
        boxes = self._chunk_grid(shape, cdims)
        cells = int(np.prod(cdims, dtype=np.int64))
        base, buf = self._base, self._buf

        def one(item):
            addr, size, mask, offs = item
            raw = self._unfilter(bytes(buf[base + addr:base + addr + size]), filters, mask)
            chunk = np.frombuffer(raw, t.dtype, count=cells).reshape(cdims)
            dst, src = boxes(offs)
            out[dst] = chunk[src]

        items = list(self._btree_leaves(btree, group=False, ndims=rank))
        if len(items) > 1 and filters and self._threads > 1:
            with ThreadPoolExecutor(self._threads) as ex:
                list(ex.map(one, items))
        else:
            for it in items:
                one(it)
        return out

    def _fixed_array_chunks(self, ds: Dataset, addr: int, cdims: Sequence[int], shape) -> np.ndarray:
        """Chunks indexed by a fixed array (header FAHD, data block FADB), one
        entry per chunk in row-major order of the chunk grid; entries of more
        than 2^page_bits chunks are paged, pages marked in a bitmap (most
        significant bit first) and each followed by its checksum."""
        t = ds._type
        out = np.zeros(shape, t.dtype)
        if addr == UNDEF:
            return out
        c = self._cur(addr)
        if c.raw(4) != b"FAHD":
            raise HDF5Error(f"{ds.name}: fixed array header signature")
        c.u(1)
        client, esz, pbits = c.u(1), c.u(1), c.u(1)
        n = c.length()
        dblk = c.off()
        d = self._cur(dblk)
        if d.raw(4) != b"FADB":
            raise HDF5Error(f"{ds.name}: fixed array data block signature")
        d.skip(2)
        d.off()
        per_page = 1 << pbits

        def entry(cur: _Cursor):
            a = cur.off()
            if client == 1:
                size = cur.u(esz - self._so - 4)
                return a, size, cur.u(4)
            return a, None, 0

        entries: List[Tuple[int, Optional[int], int]] = []
        if n <= per_page:
            entries = [entry(d) for _ in range(n)]
        else:
            npages = -(-n // per_page)
            bitmap = d.raw((npages + 7) // 8)
            d.skip(4)
            page0 = d.pos
            for k in range(npages):
                m = min(per_page, n - k * per_page)
                if not bitmap[k // 8] & (0x80 >> (k % 8)):
                    entries.extend([(UNDEF, None, 0)] * m)
                    continue
                pc = _Cursor(self._buf, page0 + k * (per_page * esz + 4), self._so, self._sl)
                entries.extend(entry(pc) for _ in range(m))
        filters = self._filters(ds._node)
        cells = int(np.prod(cdims, dtype=np.int64))
        nb = cells * t.size
        grid = [-(-s // n_) for s, n_ in zip(shape, cdims)]
        boxes = self._chunk_grid(shape, cdims)
        base, buf = self._base, self._buf

        def one(k_idx):
            k, idx = k_idx
            a, size, mask = entries[k]
            if a == UNDEF:
                return
            raw = bytes(buf[base + a:base + a + (size if size is not None else nb)])
            if filters:
                raw = self._unfilter(raw, filters, mask)
            chunk = np.frombuffer(raw, t.dtype, count=cells).reshape(cdims)
            dst, src = boxes(tuple(i * n_ for i, n_ in zip(idx, cdims)))
            out[dst] = chunk[src]

        work = list(enumerate(np.ndindex(*grid)))
        if len(work) > 1 and filters and self._threads > 1:
            with ThreadPoolExecutor(self._threads) as ex:
                list(ex.map(one, work))
        else:
            for w in work:
                one(w)
        return out

    def _implicit_chunks(self, ds: Dataset, addr: int, cdims: Sequence[int], shape) -> np.ndarray:
        t = ds._type
        out = np.zeros(shape, t.dtype)
        if addr == UNDEF:
            return out
        grid = [-(-s // n) for s, n in zip(shape, cdims)]
        cells = int(np.prod(cdims, dtype=np.int64))
        nb = cells * t.size
        boxes = self._chunk_grid(shape, cdims)
        p = self._base + addr
        for k, idx in enumerate(np.ndindex(*grid)):
            offs = tuple(i * n for i, n in zip(idx, cdims))
            chunk = np.frombuffer(bytes(self._buf[p + k * nb:p + (k + 1) * nb]), t.dtype).reshape(cdims)
            dst, src = boxes(offs)
            out[dst] = chunk[src]
        return out


# ----------------------------------------------------------------------------- writer
class _Writer:
    """Version-0 files (HDF5's earliest format, what h5py writes by default):
    groups as symbol tables, chunked datasets under version-1 B-trees with
    the deflate filter, contiguous datasets, scalar strings and numbers."""

    def __init__(self, threads: int = 1):
        self.buf = bytearray()
        self.threads = threads

    def alloc(self, n: int) -> int:
        a = len(self.buf)
        self.buf += bytes(_align8(n))
        return a

    def put(self, addr: int, data: bytes) -> None:
        self.buf[addr:addr + len(data)] = data

    def append(self, data: bytes) -> int:
        a = self.alloc(len(data))
        self.put(a, data)
        return a

    # messages
    @staticmethod
    def msg(mtype: int, data: bytes, flags: int = 0) -> bytes:
        data = data + bytes(_align8(len(data)) - len(data))
        return struct.pack("<HHB3x", mtype, len(data), flags) + data

    def header(self, msgs: List[bytes]) -> int:
        """A version-1 object header; a null message pads it to 256 bytes of
        message space, as the library sizes new headers."""
        body = b"".join(msgs)
        n = len(msgs)
        size = max(len(body), 256)
        if size > len(body):
            pad = size - len(body)
            body += struct.pack("<HHB3x", MSG_NIL, pad - 8, 0) + bytes(pad - 8)
            n += 1
        return self.append(struct.pack("<BBHII", 1, 0, n, 1, size) + bytes(4) + body)

    @staticmethod
    def dataspace(shape: Tuple[int, ...]) -> bytes:
        return struct.pack("<BBB5x", 1, len(shape), 0) + b"".join(struct.pack("<Q", int(s)) for s in shape)

    @staticmethod
    def datatype(dt: np.dtype, vstring: bool = False) -> bytes:
        if vstring:   # string of UTF-8 bytes; base type unsigned char
            base = struct.pack("<B3sIHH", 0x10 | 0, bytes([0, 0, 0]), 1, 0, 8)
            return struct.pack("<B3sI", 0x10 | 9, bytes([0x01, 0x01, 0]), 16) + base
        if dt.kind == "f":
            bits = 8 * dt.itemsize
            if dt.itemsize == 4:
                props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
            else:
                props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
            return struct.pack("<B3sI", 0x10 | 1, bytes([0x20, bits - 1, 0]), dt.itemsize) + props
        if dt.kind in "iub":
            signed = dt.kind == "i"
            return struct.pack("<B3sIHH", 0x10 | 0, bytes([0x08 if signed else 0, 0, 0]), dt.itemsize, 0,
                               8 * dt.itemsize)
        if dt.kind == "S":
            return struct.pack("<B3sI", 0x10 | 3, bytes([0, 0, 0]), dt.itemsize)
        raise NotImplementedError(f"writing dtype {dt}")

    def layout_contiguous(self, addr: int, nbytes: int) -> bytes:
        return struct.pack("<BBQQ", 3, 1, addr, nbytes)

    def layout_chunked(self, btree: int, cdims: Tuple[int, ...], esize: int) -> bytes:
        return struct.pack("<BBB", 3, 2, len(cdims) + 1) + struct.pack("<Q", btree) + \
            b"".join(struct.pack("<I", d) for d in cdims) + struct.pack("<I", esize)

    @staticmethod
    def filters_deflate(level: int) -> bytes:
        name = b"deflate\0"
        return struct.pack("<BB6x", 1, 1) + struct.pack("<HHHH", FILTER_DEFLATE, len(name), 1, 1) + name + \
            struct.pack("<I", level) + bytes(4)

    def chunk_btree(self, chunks: List[Tuple[Tuple[int, ...], int, int]], rank: int,
                    cdims: Tuple[int, ...]) -> int:
        """Nodes of up to 2K = 64 entries (the default indexed-storage K of a
        version-0 superblock), parent levels as needed.  chunks: (offsets,
        address, stored size), in offset order.  A node's last key is the
        corner past its last chunk, so that it scales above that chunk."""
        K = 32

        def node(level: int, entries, last_key) -> int:
            n = len(entries)
            body = struct.pack("<4sBBHQQ", b"TREE", 1, level, n, UNDEF, UNDEF)
            for (key, child) in entries:
                body += key + struct.pack("<Q", child)
            body += last_key
            body += bytes((2 * K - n) * (8 + 8 + 8 * (rank + 1)))   # room for a full node, as the library reserves
            return self.append(body)

        def key(size: int, offs: Tuple[int, ...]) -> bytes:
            return struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in offs) + struct.pack("<Q", 0)

        leaves = []
        for i in range(0, len(chunks), 2 * K):
            part = chunks[i:i + 2 * K]
            ents = [(key(sz, offs), a) for offs, a, sz in part]
            nxt = chunks[i + 2 * K][0] if i + 2 * K < len(chunks) else None
            last = key(0, nxt) if nxt is not None else key(0, tuple(o + c for o, c in zip(part[-1][0], cdims)))
            leaves.append((ents[0][0], node(0, ents, last), last))
        if len(leaves) == 1:
            return leaves[0][1]
        level = 1
        while len(leaves) > 1:
            parents = []
            for i in range(0, len(leaves), 2 * K):
                part = leaves[i:i + 2 * K]
                parents.append((part[0][0], node(level, [(k0, a) for k0, a, _ in part], part[-1][2]), part[-1][2]))
            leaves = parents
            level += 1
        return leaves[0][1]

    def global_heap(self, strings: List[bytes]) -> Tuple[int, List[int]]:
        objs = b""
        for k, s in enumerate(strings, start=1):
            objs += struct.pack("<HH4xQ", k, 1, len(s)) + s + bytes(_align8(len(s)) - len(s))
        size = max(4096, 16 + len(objs) + 16)
        body = struct.pack("<4sB3xQ", b"GCOL", 1, size) + objs
        free = size - len(body)
        body += struct.pack("<HH4xQ", 0, 0, free) + bytes(free - 16)
        return self.append(body), list(range(1, len(strings) + 1))

    def group(self, members: Dict[str, int]) -> int:
        names = sorted(members)
        heap_data = bytearray(b"\0" * 8)
        offs = {}
        for nm in names:
            offs[nm] = len(heap_data)
            b = nm.encode("utf-8") + b"\0"
            heap_data += b + bytes(_align8(len(b)) - len(b))
        heap_size = max(len(heap_data) + 16, 88)
        free_off = len(heap_data)
        heap_data += bytes(heap_size - len(heap_data))
        struct.pack_into("<QQ", heap_data, free_off, 1, heap_size - free_off)   # free block: next = 1 (none), size
        data_addr = self.append(bytes(heap_data))
        heap = self.append(struct.pack("<4sB3xQQQ", b"HEAP", 0, heap_size, free_off, data_addr))
        # symbol nodes of up to 2K = 8 entries, under one version-1 B-tree node
        snods = []
        for i in range(0, max(len(names), 1), 8):
            part = names[i:i + 8]
            body = struct.pack("<4sBBH", b"SNOD", 1, 0, len(part))
            for nm in part:
                body += struct.pack("<QQII16x", offs[nm], members[nm], 0, 0)
            body += bytes((8 - len(part)) * 40)
            snods.append((part, self.append(body)))
        body = struct.pack("<4sBBHQQ", b"TREE", 0, 0, len(snods), UNDEF, UNDEF)
        body += struct.pack("<Q", 0)
        for part, a in snods:
            body += struct.pack("<Q", a) + struct.pack("<Q", offs[part[-1]] if part else 0)
        body += bytes((32 - len(snods)) * 16)
        btree = self.append(body)
        return self.header([self.msg(MSG_SYMBOL_TABLE, struct.pack("<QQ", btree, heap))])

    def dataset(self, value, chunks: Optional[Tuple[int, ...]] = None, level: int = 4) -> int:
        if isinstance(value, str):
            value = value.encode("utf-8")
            coll, idx = self.global_heap([value])
            raw = struct.pack("<IQI", len(value), coll, idx[0])
            data = self.append(raw)
            return self.header([self.msg(MSG_DATASPACE, self.dataspace(())),
                                self.msg(MSG_DATATYPE, self.datatype(np.dtype(object), vstring=True), 1),
                                self.msg(MSG_LAYOUT, self.layout_contiguous(data, len(raw)))])
        a = np.asarray(value)
        if not a.flags.c_contiguous:
            a = a.copy(order="C")   # (ascontiguousarray would make a 0-d value 1-d)
        if a.dtype == np.bool_:
            a = a.astype(np.int8)
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        shape = a.shape
        msgs = [self.msg(MSG_DATASPACE, self.dataspace(shape)), self.msg(MSG_DATATYPE, self.datatype(a.dtype), 1)]
        if chunks is None:
            addr = self.append(a.tobytes()) if a.nbytes else UNDEF
            msgs.append(self.msg(MSG_LAYOUT, self.layout_contiguous(addr, a.nbytes)))
            return self.header(msgs)
        grid = [-(-s // n) for s, n in zip(shape, chunks)]
        offsets = [tuple(i * n for i, n in zip(idx, chunks)) for idx in np.ndindex(*grid)]

        def compress(offs):
            block = np.zeros(chunks, a.dtype)
            src = tuple(slice(o, min(o + n, s)) for o, n, s in zip(offs, chunks, shape))
            block[tuple(slice(0, sl.stop - sl.start) for sl in src)] = a[src]
            return zlib.compress(block.tobytes(), level)

        if len(offsets) > 1 and self.threads > 1:   # zlib releases the GIL
            with ThreadPoolExecutor(self.threads) as ex:
                comps = list(ex.map(compress, offsets))
        else:
            comps = [compress(o) for o in offsets]
        entries = [(offs, self.append(comp), len(comp)) for offs, comp in zip(offsets, comps)]
        btree = self.chunk_btree(entries, len(shape), tuple(chunks))
        msgs.append(self.msg(MSG_FILTERS, self.filters_deflate(level)))
        msgs.append(self.msg(MSG_LAYOUT, self.layout_chunked(btree, tuple(chunks), a.dtype.itemsize)))
        return self.header(msgs)


def write(path: str, tree: Dict[str, object], chunked: Optional[Dict[str, Tuple[int, ...]]] = None,
          level: int = 4, threads: Optional[int] = None) -> None:
    """Write a nested dict as an HDF5 file in the reference's layout.

    tree: name -> numpy array / number / str (a dataset) or dict (a group).
    chunked: dataset path -> chunk shape, stored with gzip `level` (the
    reference stores ux, uy in chunks of one frame, gzip 4, PF:908-923)."""
    chunked = chunked or {}
    w = _Writer(threads or min(16, os.cpu_count() or 1))
    sb_size = 8 + 16 + 4 * 8 + 40
    w.alloc(sb_size)

    def build(node: Dict[str, object], prefix: str) -> int:
        members = {}
        for k, v in node.items():
            p = f"{prefix}/{k}" if prefix else k
            if isinstance(v, dict):
                members[k] = build(v, p)
            else:
                members[k] = w.dataset(v, chunked.get(p), level)
        return w.group(members)

    root = build(tree, "")
    # the root group's symbol table (B-tree, heap: its header's first
    # message) is repeated in the superblock's root entry scratch pad
    mtype, = struct.unpack_from("<H", w.buf, root + 16)
    assert mtype == MSG_SYMBOL_TABLE
    st = struct.unpack_from("<QQ", w.buf, root + 24)
    sb = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", 4, 16, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, len(w.buf), UNDEF)
    sb += struct.pack("<QQII", 0, root, 1, 0) + struct.pack("<QQ", *st)
    w.put(0, sb)
    with open(path, "wb") as fh:
        fh.write(bytes(w.buf))
