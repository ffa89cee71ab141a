"""A small HDF5 reader (numpy only) for the files rubix writes.

rubix stores galaxies as rubix_galaxy.h5 through h5py
(rubix/galaxy/input_handler/base.py:16-75) and reads them back with
h5py in load_galaxy_data (rubix/utils.py:153-197).  h5py is not available
here, so this module reads the HDF5 file format directly (HDF5 File Format
Specification, version 3): it supports what h5py writes for such files --
superblocks 0-3, version 1 and 2 object headers (with continuation blocks),
old-style groups (symbol table: v1 B-tree + symbol nodes + local heap) and
compact new-style groups (link messages), contiguous, compact and chunked
(v1 B-tree index, deflate and shuffle filters) datasets, fixed-point,
floating-point, fixed-length string and variable-length string data, and
attributes.  Dense link or attribute storage (fractal heaps), the version 4
chunk indexes, compound and reference types are rejected with a clear error.

    with H5File(path) as f:
        x = f["particles/stars/coords"][()]
        unit = f["particles/stars/coords"].attrs["unit"]
        names = f["particles"].keys()
"""
from __future__ import annotations

import zlib
from typing import Dict, List

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEFINED = 0xFFFFFFFFFFFFFFFF


class HDF5Error(ValueError):
    pass


class _Buf:
    """Little-endian cursor over bytes."""

    def __init__(self, data: bytes, pos: int = 0):
        self.data = data
        self.pos = pos

    def take(self, n: int) -> bytes:
        if self.pos + n > len(self.data):
            raise HDF5Error("truncated HDF5 structure")
        b = self.data[self.pos:self.pos + n]
        self.pos += n
        return b

    def u(self, n: int) -> int:
        return int.from_bytes(self.take(n), "little")

    def skip(self, n: int) -> None:
        self.take(n)

    def align(self, base: int, a: int = 8) -> None:
        self.pos = base + (self.pos - base + a - 1) // a * a


# ------------------------------------------------------------------ datatypes

class _Type:
    """A decoded datatype message: numpy dtype for fixed-size data, or a
    variable-length string marker."""

    def __init__(self, cls: int, size: int, dtype=None, vlen_str=False, encoding="ascii"):
        self.cls = cls
        self.size = size
        self.dtype = dtype
        self.vlen_str = vlen_str
        self.encoding = encoding


def _parse_type(b: _Buf) -> _Type:
    cv = b.u(1)
    cls, _version = cv & 0x0F, cv >> 4
    bits = b.u(3)
    size = b.u(4)
    if cls == 0:  # fixed-point
        order = ">" if bits & 1 else "<"
        signed = "i" if bits & 0x08 else "u"
        b.skip(4)  # bit offset, precision
        return _Type(cls, size, np.dtype(f"{order}{signed}{size}"))
    if cls == 1:  # floating-point
        order = ">" if bits & 1 else "<"
        if bits & 0x40:
            raise HDF5Error("VAX floating-point order is not supported")
        b.skip(12)  # offsets, exponent / mantissa layout, bias
        if size not in (2, 4, 8):
            raise HDF5Error(f"unsupported float size {size}")
        return _Type(cls, size, np.dtype(f"{order}f{size}"))
    if cls == 3:  # fixed-length string
        enc = "utf-8" if (bits >> 4) & 0x0F == 1 else "ascii"
        return _Type(cls, size, np.dtype(f"S{size}"), encoding=enc)
    if cls == 9:  # variable-length
        vtype = bits & 0x0F
        enc = "utf-8" if (bits >> 8) & 0x0F == 1 else "ascii"
        _parse_type(b)  # base type (a character type for strings)
        if vtype != 1:
            raise HDF5Error("variable-length sequences are not supported (only strings)")
        return _Type(cls, size, None, vlen_str=True, encoding=enc)
    if cls == 8:  # enumeration (h5py booleans): read as the base integer type
        base = _parse_type(b)
        return _Type(cls, size, base.dtype)
    raise HDF5Error(f"unsupported HDF5 datatype class {cls}")


def _parse_space(b: _Buf, size_len: int):
    version = b.u(1)
    rank = b.u(1)
    flags = b.u(1)
    if version == 1:
        b.skip(5)
        kind = 1 if rank > 0 else 0
    elif version == 2:
        kind = b.u(1)  # 0 scalar, 1 simple, 2 null
    else:
        raise HDF5Error(f"unsupported dataspace version {version}")
    dims = tuple(b.u(size_len) for _ in range(rank))
    if flags & 1:
        b.skip(size_len * rank)
    if kind == 2:
        return None  # null dataspace: no elements
    return dims


# ------------------------------------------------------------------ file

class H5File:
    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.data = fh.read()
        self.path = path
        base = 0
        while self.data[base:base + 8] != SIGNATURE:
            base = 512 if base == 0 else base * 2
            if base + 8 > len(self.data):
                raise HDF5Error(f"{path}: not an HDF5 file")
        b = _Buf(self.data, base + 8)
        version = b.u(1)
        if version in (0, 1):
            b.skip(3)  # free-space, root symbol table, reserved
            b.skip(1)  # shared header version
            self.so = b.u(1)
            self.sl = b.u(1)
            b.skip(1)
            b.skip(4)  # group K values
            b.skip(4)  # consistency flags
            if version == 1:
                b.skip(4)
            self.base = b.u(self.so)
            b.skip(self.so * 3)  # free space, end of file, driver info
            b.skip(self.so)  # root entry: link name offset
            root = b.u(self.so)
        elif version in (2, 3):
            self.so = b.u(1)
            self.sl = b.u(1)
            b.skip(1)
            self.base = b.u(self.so)
            b.skip(self.so * 2)  # superblock extension, end of file
            root = b.u(self.so)
        else:
            raise HDF5Error(f"unsupported superblock version {version}")
        if self.base:  # a user block precedes the data: addresses are relative
            self.data = self.data[self.base:]
        self.root = Group(self, root, "/")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        self.data = b""

    def __getitem__(self, path: str):
        return self.root[path]

    def keys(self):
        return self.root.keys()

    # -- low-level helpers
    def buf(self, addr: int) -> _Buf:
        if addr == UNDEFINED or addr < 0 or addr >= len(self.data):
            raise HDF5Error(f"address {addr:#x} outside the file")
        return _Buf(self.data, addr)

    def messages(self, addr: int) -> List[tuple]:
        """(type, flags, raw bytes) of every message of the object header at addr."""
        b = self.buf(addr)
        out = []
        if self.data[addr:addr + 4] == b"OHDR":
            b.skip(4)
            version = b.u(1)
            if version != 2:
                raise HDF5Error(f"unsupported object header version {version}")
            flags = b.u(1)
            if flags & 0x20:
                b.skip(16)
            if flags & 0x10:
                b.skip(4)
            size = b.u(1 << (flags & 3))
            self._v2_messages(b, b.pos + size, flags, out)
            return out
        version = b.u(1)
        if version != 1:
            raise HDF5Error(f"unsupported object header version {version}")
        b.skip(1)
        nmsg = b.u(2)
        b.skip(4)  # reference count
        size = b.u(4)
        b.skip(4)  # alignment of the prefix to 8 bytes
        blocks = [(b.pos, size)]
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            c = _Buf(self.data, start)
            while c.pos + 8 <= start + length and len(out) < nmsg:
                mtype = c.u(2)
                msize = c.u(2)
                mflags = c.u(1)
                c.skip(3)
                raw = c.take(msize)
                if mtype == 0x10:
                    cb = _Buf(raw)
                    blocks.append((cb.u(self.so), cb.u(self.sl)))
                out.append((mtype, mflags, raw))
        return out

    def _v2_messages(self, b: _Buf, end: int, flags: int, out: list) -> None:
        """Messages of one version 2 header block ending at `end` (the
        block's checksum follows it); a gap shorter than a message header
        may close the block."""
        hdr = 6 if flags & 0x04 else 4
        conts = []
        while b.pos + hdr <= end:
            mtype = b.u(1)
            msize = b.u(2)
            mflags = b.u(1)
            if flags & 0x04:
                b.skip(2)  # creation order
            if b.pos + msize > end:
                raise HDF5Error("object header message runs past its block")
            raw = b.take(msize)
            if mtype == 0x10:
                cb = _Buf(raw)
                conts.append((cb.u(self.so), cb.u(self.sl)))
            if mtype != 0:
                out.append((mtype, mflags, raw))
        for addr, length in conts:
            if self.data[addr:addr + 4] != b"OCHK":
                raise HDF5Error("bad object header continuation block")
            self._v2_messages(_Buf(self.data, addr + 4), addr + length - 4, flags, out)

    def local_heap_name(self, heap_addr: int, offset: int) -> str:
        b = self.buf(heap_addr)
        if b.take(4) != b"HEAP":
            raise HDF5Error("bad local heap")
        b.skip(4)
        b.skip(self.sl * 2)
        data_addr = b.u(self.so)
        start = data_addr + offset
        end = self.data.index(b"\0", start)
        return self.data[start:end].decode("utf-8")

    def global_heap_object(self, coll_addr: int, index: int) -> bytes:
        b = self.buf(coll_addr)
        if b.take(4) != b"GCOL":
            raise HDF5Error("bad global heap collection")
        b.skip(4)
        csize = b.u(self.sl)
        end = coll_addr + csize
        while b.pos + 8 + self.sl <= end:
            idx = b.u(2)
            b.skip(6)
            osize = b.u(self.sl)
            if idx == 0:
                break
            data = b.take(osize)
            if idx == index:
                return data
            b.align(coll_addr, 8)
        raise HDF5Error(f"global heap object {index} not found")


# ------------------------------------------------------------------ objects

class _Object:
    def __init__(self, f: H5File, addr: int, name: str):
        self.file = f
        self.addr = addr
        self.name = name
        self._msgs = f.messages(addr)

    def _msg(self, mtype: int):
        return [raw for t, _, raw in self._msgs if t == mtype]

    @property
    def attrs(self) -> Dict[str, object]:
        out = {}
        for t, _, raw in self._msgs:
            if t == 0x000C:
                name, value = _parse_attribute(self.file, raw)
                out[name] = value
            elif t == 0x0015:
                b = _Buf(raw)
                b.skip(1)
                fl = b.u(1)
                if fl & 1:
                    b.skip(2)
                if b.u(self.file.so) != UNDEFINED:
                    raise HDF5Error(f"{self.name}: dense attribute storage is not supported")
        return out


class Group(_Object):
    def _links(self) -> Dict[str, int]:
        f = self.file
        links: Dict[str, int] = {}
        for raw in self._msg(0x0011):  # symbol table: v1 B-tree of symbol nodes
            b = _Buf(raw)
            btree, heap = b.u(f.so), b.u(f.so)
            self._walk_group_btree(btree, heap, links)
        for raw in self._msg(0x0006):  # link messages (compact new-style group)
            name, addr = _parse_link(f, raw)
            if addr is not None:
                links[name] = addr
        for raw in self._msg(0x0002):  # link info: dense storage
            b = _Buf(raw)
            b.skip(1)
            fl = b.u(1)
            if fl & 1:
                b.skip(8)
            heap = b.u(f.so)
            name_index = b.u(f.so)
            if heap != UNDEFINED:
                for obj in _dense_objects(f, heap, name_index, self.name):
                    name, addr = _parse_link(f, obj)
                    if addr is not None:
                        links[name] = addr
        return links

    def _walk_group_btree(self, addr: int, heap: int, links: Dict[str, int]) -> None:
        f = self.file
        b = f.buf(addr)
        if b.take(4) != b"TREE":
            raise HDF5Error("bad group B-tree node")
        if b.u(1) != 0:
            raise HDF5Error("group B-tree of the wrong node type")
        level = b.u(1)
        used = b.u(2)
        b.skip(2 * f.so)  # siblings
        children = []
        for _ in range(used):
            b.skip(f.sl)  # key
            children.append(b.u(f.so))
        for child in children:
            if level > 0:
                self._walk_group_btree(child, heap, links)
                continue
            s = f.buf(child)
            if s.take(4) != b"SNOD":
                raise HDF5Error("bad symbol table node")
            s.skip(2)
            n = s.u(2)
            for _ in range(n):
                name_off = s.u(f.so)
                obj = s.u(f.so)
                s.skip(24)  # cache type, reserved, scratch pad
                links[f.local_heap_name(heap, name_off)] = obj

    def keys(self) -> List[str]:
        return sorted(self._links())

    def __contains__(self, name: str) -> bool:
        try:
            self[name]
            return True
        except KeyError:
            return False

    def __getitem__(self, path: str):
        parts = [p for p in path.split("/") if p]
        obj = self
        for p in parts:
            if not isinstance(obj, Group):
                raise KeyError(path)
            links = obj._links()
            if p not in links:
                raise KeyError(f"{path!r} not found in {self.file.path}")
            addr = links[p]
            full = (obj.name.rstrip("/") + "/" + p)
            msgs = [t for t, _, _ in self.file.messages(addr)]
            obj = Group(self.file, addr, full) if (0x0011 in msgs or 0x0006 in msgs or 0x0002 in msgs) \
                and 0x0008 not in msgs else Dataset(self.file, addr, full)
        return obj


class Dataset(_Object):
    @property
    def shape(self):
        raw = self._msg(0x0001)
        if not raw:
            raise HDF5Error(f"{self.name}: no dataspace")
        return _parse_space(_Buf(raw[0]), self.file.sl)

    @property
    def _type(self) -> _Type:
        for t, flags, raw in self._msgs:
            if t == 0x0003:
                if flags & 0x02:  # the message holds a reference to a committed type
                    raise HDF5Error(f"{self.name}: committed (shared) datatypes are not supported")
                return _parse_type(_Buf(raw))
        raise HDF5Error(f"{self.name}: no datatype")

    def __getitem__(self, key):
        if key != () and key != Ellipsis:
            return self[()][key]
        return self.read()

    def read(self):
        f = self.file
        shape = self.shape
        t = self._type
        if shape is None:
            return np.empty((0,), dtype=t.dtype if t.dtype is not None else object)
        count = int(np.prod(shape)) if shape else 1
        raw = self._raw_bytes(count * t.size, shape, t.size)
        arr = _decode(f, raw, t, count, as_str=False)
        return arr.reshape(shape) if shape else arr.reshape(())[()]

    def _raw_bytes(self, nbytes: int, shape, esize: int) -> bytes:
        f = self.file
        raw = self._msg(0x0008)
        if not raw:
            raise HDF5Error(f"{self.name}: no data layout")
        b = _Buf(raw[0])
        version = b.u(1)
        if version in (1, 2):
            rank = b.u(1)
            cls = b.u(1)
            b.skip(5)
            addr = b.u(f.so) if cls > 0 else None
            dims = [b.u(4) for _ in range(rank)]
            if cls == 1:
                return f.data[addr:addr + nbytes] if addr != UNDEFINED else bytes(nbytes)
            if cls == 0:
                b.skip(4)
                return b.take(nbytes)
            return self._chunked(addr, dims[:-1] if len(dims) > len(shape) else dims, shape, esize)
        if version not in (3, 4):
            raise HDF5Error(f"{self.name}: unsupported layout version {version}")
        cls = b.u(1)
        if cls == 0:  # compact
            n = b.u(2)
            return b.take(n)[:nbytes]
        if cls == 1:  # contiguous
            addr = b.u(f.so)
            if addr == UNDEFINED:
                return bytes(nbytes)  # never written: fill value 0
            return f.data[addr:addr + nbytes]
        if cls == 2 and version == 3:
            rank = b.u(1)
            addr = b.u(f.so)
            dims = [b.u(4) for _ in range(rank)]
            return self._chunked(addr, dims[:-1], shape, esize)
        raise HDF5Error(f"{self.name}: chunk index of layout version {version} is not supported")

    def _filters(self):
        out = []
        for raw in self._msg(0x000B):
            b = _Buf(raw)
            version = b.u(1)
            n = b.u(1)
            if version == 1:
                b.skip(6)
            for _ in range(n):
                fid = b.u(2)
                name_len = b.u(2) if (version == 1 or fid >= 256) else 0
                b.skip(2)  # flags
                nvals = b.u(2)
                b.skip(name_len)  # version 1 pads the name to a multiple of 8
                vals = [b.u(4) for _ in range(nvals)]
                if version == 1 and nvals % 2:
                    b.skip(4)
                out.append((fid, vals))
        return out

    def _chunked(self, btree: int, cdims, shape, esize: int) -> bytes:
        f = self.file
        rank = len(shape)
        filters = self._filters()
        out = np.zeros(int(np.prod(shape)) * esize, dtype=np.uint8)
        dst = out.reshape(tuple(shape) + (esize,))
        if btree == UNDEFINED:
            return out.tobytes()

        def walk(addr):
            b = f.buf(addr)
            if b.take(4) != b"TREE":
                raise HDF5Error("bad chunk B-tree node")
            if b.u(1) != 1:
                raise HDF5Error("chunk B-tree of the wrong node type")
            level = b.u(1)
            used = b.u(2)
            b.skip(2 * f.so)
            for _ in range(used):
                csize = b.u(4)
                mask = b.u(4)
                offs = [b.u(8) for _ in range(rank + 1)][:rank]
                child = b.u(f.so)
                if level > 0:
                    walk(child)
                    continue
                data = f.data[child:child + csize]
                for i, (fid, vals) in reversed(list(enumerate(filters))):
                    if mask & (1 << i):
                        continue
                    if fid == 1:
                        data = zlib.decompress(data)
                    elif fid == 2:
                        arr = np.frombuffer(data, dtype=np.uint8)
                        data = arr.reshape(esize, -1).T.copy().tobytes()
                    elif fid == 3:
                        data = data[:-4]  # Fletcher32 checksum
                    else:
                        raise HDF5Error(f"{self.name}: unsupported filter {fid}")
                chunk = np.frombuffer(data, dtype=np.uint8).reshape(tuple(cdims) + (esize,))
                sl_dst, sl_src = [], []
                for d in range(rank):
                    lo = offs[d]
                    hi = min(lo + cdims[d], shape[d])
                    sl_dst.append(slice(lo, hi))
                    sl_src.append(slice(0, hi - lo))
                dst[tuple(sl_dst)] = chunk[tuple(sl_src)]
            b.skip(8 * (rank + 1) + 8)  # final key

        walk(btree)
        return out.tobytes()


def _decode(f: H5File, raw: bytes, t: _Type, count: int, as_str: bool = True):
    """Element values; strings as str (attributes, as h5py returns them) or
    as bytes (dataset reads, as h5py 3 returns them)."""
    if t.vlen_str:
        out = np.empty(count, dtype=object)
        b = _Buf(raw)
        for i in range(count):
            n = b.u(4)
            coll = b.u(f.so)
            idx = b.u(4)
            s = f.global_heap_object(coll, idx)[:n] if n else b""
            out[i] = s.decode(t.encoding) if as_str else s
        return out
    if len(raw) < count * t.size:
        raise HDF5Error("dataset extends past the end of the file")
    arr = np.frombuffer(raw[:count * t.size], dtype=t.dtype).copy()
    if t.cls == 3:
        if as_str:
            return np.array([x.rstrip(b"\0").decode(t.encoding) for x in arr], dtype=object)
        return arr
    return arr.astype(t.dtype.newbyteorder("="), copy=False)


def _parse_attribute(f: H5File, raw: bytes):
    b = _Buf(raw)
    version = b.u(1)
    flags = b.u(1)  # reserved (v1) / flags (v2, v3)
    if version >= 2 and flags & 0x03:
        raise HDF5Error("attributes with shared datatypes or dataspaces are not supported")
    name_size, type_size, space_size = b.u(2), b.u(2), b.u(2)
    if version == 3:
        b.skip(1)  # name character set
    pad = version == 1
    name = b.take(name_size).rstrip(b"\0").decode("utf-8")
    if pad:
        b.align(0, 8)
    tpos = b.pos
    t = _parse_type(_Buf(raw, tpos))
    b.pos = tpos + type_size
    if pad:
        b.align(0, 8)
    spos = b.pos
    shape = _parse_space(_Buf(raw, spos), f.sl)
    b.pos = spos + space_size
    if pad:
        b.align(0, 8)
    if shape is None:
        return name, None
    count = int(np.prod(shape)) if shape else 1
    value = _decode(f, raw[b.pos:], t, count)
    value = value.reshape(shape) if shape else value.reshape(())[()]
    return name, value


def _dense_objects(f: H5File, heap_addr: int, btree_addr: int, where: str) -> List[bytes]:
    """The managed objects a dense link store indexes: the records of its
    name-index v2 B-tree (leaf root: hash + heap ID) resolved in its fractal
    heap (root direct block, or a root indirect block of direct blocks)."""
    def unsupported(what):
        return HDF5Error(f"{where}: dense link storage with {what} is not supported")

    h = f.buf(heap_addr)
    if h.take(4) != b"FRHP":
        raise HDF5Error("bad fractal heap header")
    h.skip(1)
    id_len = h.u(2)
    filt_len = h.u(2)
    flags = h.u(1)
    h.skip(4)  # max managed object size
    h.skip(f.sl + f.so + f.sl + f.so)  # huge objects, free space
    h.skip(f.sl * 8)  # managed / huge / tiny statistics
    width = h.u(2)
    start_size = h.u(f.sl)
    max_direct = h.u(f.sl)
    max_heap_bits = h.u(2)
    h.skip(2)  # starting rows
    root = h.u(f.so)
    rows = h.u(2)
    if filt_len:
        raise unsupported("filtered heap blocks")
    off_size = (max_heap_bits + 7) // 8
    dblock_hdr = 5 + f.so + off_size + (4 if flags & 2 else 0)

    # heap offset -> (file address of the block, heap offset of the block)
    blocks = []
    if rows == 0:
        blocks.append((root, 0, start_size))
    else:
        ib = f.buf(root)
        if ib.take(4) != b"FHIB":
            raise HDF5Error("bad fractal heap indirect block")
        ib.skip(1 + f.so + off_size)
        offset = 0
        for r in range(rows):
            size = start_size * (1 if r < 2 else 2 ** (r - 1))
            if size > max_direct:
                raise unsupported("nested indirect blocks")
            for _ in range(width):
                addr = ib.u(f.so)
                if addr != UNDEFINED:
                    blocks.append((addr, offset, size))
                offset += size

    def resolve(heap_id: bytes) -> bytes:
        if (heap_id[0] >> 4) & 3 != 0:
            raise unsupported("huge or tiny heap objects")
        len_size = id_len - 1 - off_size
        off = int.from_bytes(heap_id[1:1 + off_size], "little")
        length = int.from_bytes(heap_id[1 + off_size:1 + off_size + len_size], "little")
        for addr, boff, size in blocks:
            if boff <= off < boff + size:
                if off - boff < dblock_hdr:
                    break
                start = addr + (off - boff)
                return f.data[start:start + length]
        raise HDF5Error("fractal heap object outside the heap's blocks")

    t = f.buf(btree_addr)
    if t.take(4) != b"BTHD":
        raise HDF5Error("bad v2 B-tree header")
    t.skip(1)
    if t.u(1) != 5:
        raise HDF5Error("link name index of the wrong B-tree type")
    t.skip(4)  # node size
    rec_size = t.u(2)
    depth = t.u(2)
    t.skip(2)
    root_node = t.u(f.so)
    nrec = t.u(2)
    if depth != 0:
        raise unsupported("a multi-level name index")
    out = []
    if nrec == 0:
        return out
    leaf = f.buf(root_node)
    if leaf.take(4) != b"BTLF":
        raise HDF5Error("bad v2 B-tree leaf")
    leaf.skip(2)
    for _ in range(nrec):
        rec = leaf.take(rec_size)
        out.append(resolve(rec[4:4 + id_len]))  # hash, heap ID
    return out


def _parse_link(f: H5File, raw: bytes):
    b = _Buf(raw)
    b.skip(1)  # version
    flags = b.u(1)
    ltype = b.u(1) if flags & 0x08 else 0
    if flags & 0x04:
        b.skip(8)
    if flags & 0x10:
        b.skip(1)
    n = b.u(1 << (flags & 3))
    name = b.take(n).decode("utf-8")
    if ltype == 0:
        return name, b.u(f.so)
    return name, None  # soft / external links are not followed
