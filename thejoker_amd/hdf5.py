"""HDF5 tables of samples, read and written without h5py.

The reference stores prior and posterior samples as an astropy table in
HDF5 (samples.py:480-545, samples_helpers.py:35-272): one compound dataset
``samples`` (a float64 field per column, chunked, maxshape (None,), no
filters) and a dataset ``samples.__table_column_meta__`` holding the table's
YAML header, one line per element.  Its batch readers open that file with
PyTables and h5py (utils.py:168-228).

This module implements the subset of the HDF5 file format such files use,
in numpy only, so that a prior cache written by the reference streams
straight into the device columns:

* reading: superblock versions 0-3; object headers versions 1 and 2
  (with continuation blocks); groups held in a symbol table (version 1
  B-tree + local heap) or in compact link messages; dataspaces; fixed-point,
  floating-point, string, compound, array and variable-length datatypes;
  compact, contiguous and chunked layouts, chunked ones indexed by a
  version 1 B-tree or (layout message version 4) a single chunk or implicit
  index; variable-length strings through the global heap.  Filtered
  (compressed) data, dense link storage and the fixed-array, extensible-array
  and version 2 B-tree chunk indices raise NotImplementedError.
* writing: the layout the reference's writer produces with h5py's default
  file format -- superblock version 0, a root symbol table, the table as a
  chunked compound dataset with unlimited maxshape (so h5py can append to
  it, as samples_helpers.py:239-243 does) and the header as fixed-length
  strings.
"""
import mmap
import os
import struct

import numpy as np

__all__ = ["File", "Dataset", "write_table", "H5FormatError"]

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = (1 << 64) - 1


class H5FormatError(ValueError):
    """The file is not HDF5, or is damaged."""


def _uint(buf, pos, n):
    return int.from_bytes(buf[pos:pos + n], "little")


def _cstring(buf, pos):
    end = buf.find(b"\x00", pos)
    if end < 0:
        raise H5FormatError("unterminated string")
    return bytes(buf[pos:end]).decode("utf-8"), end


# ---------------------------------------------------------------------------
# datatypes
def _parse_dtype(buf, pos):
    """(numpy dtype, vlen info or None, position after the message).

    vlen info is ("str", None) for variable-length strings and
    ("seq", base dtype) for sequences; their elements are read as the raw
    16-byte (length, collection address, object index) records."""
    cv = buf[pos]
    cls, ver = cv & 0x0F, cv >> 4
    bits = buf[pos + 1] | (buf[pos + 2] << 8) | (buf[pos + 3] << 16)
    size = _uint(buf, pos + 4, 4)
    p = pos + 8
    if cls == 0:  # fixed point
        order = ">" if bits & 1 else "<"
        kind = "i" if bits & 0x08 else "u"
        return np.dtype(f"{order}{kind}{size}"), None, p + 4
    if cls == 1:  # floating point
        if bits & 0x40:
            raise NotImplementedError("VAX floating point")
        order = ">" if bits & 1 else "<"
        return np.dtype(f"{order}f{size}"), None, p + 12
    if cls == 3:  # fixed-length string
        return np.dtype(f"S{size}"), None, p
    if cls == 4:  # bit field
        return np.dtype(f"{'>' if bits & 1 else '<'}u{size}"), None, p + 4
    if cls == 5:  # opaque: tag string padded to a multiple of 8
        tag_len = bits & 0xFF
        return np.dtype(f"V{size}"), None, p + tag_len
    if cls == 6:  # compound
        nmemb = bits & 0xFFFF
        names, formats, offsets = [], [], []
        for _ in range(nmemb):
            name, end = _cstring(buf, p)
            if ver < 3:
                p = p + ((end + 1 - p + 7) // 8) * 8
                off = _uint(buf, p, 4)
                p += 4
                shape = ()
                if ver == 1:
                    ndim = buf[p]
                    dims = [_uint(buf, p + 12 + 4 * k, 4) for k in range(4)]
                    shape = tuple(dims[:ndim])
                    p += 28
            else:
                p = end + 1
                width = 1 if size < 1 << 8 else 2 if size < 1 << 16 else 3 if size < 1 << 24 else 4
                off = _uint(buf, p, width)
                p += width
                shape = ()
            sub, vlen, p = _parse_dtype(buf, p)
            if vlen is not None:
                raise NotImplementedError("variable-length compound members")
            names.append(name)
            formats.append((sub, shape) if shape else sub)
            offsets.append(off)
        dt = np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})
        return dt, None, p
    if cls == 9:  # variable length
        base, _, p2 = _parse_dtype(buf, p)
        info = ("str", None) if (bits & 0x0F) == 1 else ("seq", base)
        return np.dtype("V%d" % size), info, p2
    if cls == 10:  # array
        ndim = buf[p]
        if ver < 3:
            dims = [_uint(buf, p + 4 + 4 * k, 4) for k in range(ndim)]
            p = p + 4 + 8 * ndim
        else:
            dims = [_uint(buf, p + 1 + 4 * k, 4) for k in range(ndim)]
            p = p + 1 + 4 * ndim
        base, vlen, p = _parse_dtype(buf, p)
        if vlen is not None:
            raise NotImplementedError("arrays of variable-length elements")
        return np.dtype((base, tuple(dims))), None, p
    raise NotImplementedError(f"HDF5 datatype class {cls}")


# ---------------------------------------------------------------------------
class Dataset:
    """One dataset: shape, maxshape, numpy dtype and its storage.  Data are
    read lazily from the memory-mapped file (:meth:`read`)."""

    def __init__(self, f, name, messages):
        self.file, self.name = f, name
        self.shape = self.maxshape = None
        self.dtype = self._vlen = None
        self._layout = None
        for mtype, data in messages:
            if mtype == 0x0001:
                self.shape, self.maxshape = f._dataspace(data)
            elif mtype == 0x0003:
                self.dtype, self._vlen, _ = _parse_dtype(data, 0)
            elif mtype == 0x0008:
                self._layout = f._layout(data)
            elif mtype == 0x000B:
                nfilters = data[1]
                if nfilters:
                    raise NotImplementedError(
                        f"dataset '{name}' uses {nfilters} HDF5 filter(s) (compression); the "
                        "sample tables are written without filters")
            elif mtype == 0x0007:
                raise NotImplementedError(f"dataset '{name}' is stored in external files")
        if self.shape is None or self.dtype is None or self._layout is None:
            raise H5FormatError(f"'{name}' is not a dataset")

    def __len__(self):
        return int(self.shape[0]) if self.shape else 1

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1

    def _raw(self, start, stop, buf=None):
        """Elements [start, stop) of the flattened dataset in the stored dtype
        (copied out of the file; into the uint8 array `buf` when given, so
        that a caller reading block after block reuses one host buffer)."""
        f, kind = self.file, self._layout[0]
        n = stop - start
        dt = self.dtype
        if buf is not None and kind in ("contiguous", "chunked") and \
                (kind == "contiguous" or len(self._layout[1]) == 1):
            if buf.dtype != np.uint8 or buf.size < n * dt.itemsize:
                raise ValueError("buf must be a uint8 array of at least the block's bytes")
            buf = buf[:n * dt.itemsize]
        else:
            buf = None
        if kind == "compact":
            return np.frombuffer(self._layout[1], dtype=dt, count=self.size)[start:stop].copy()
        if kind == "contiguous":
            addr = self._layout[1]
            if addr == UNDEF:
                return np.zeros(n, dtype=dt)
            out = np.empty(n * dt.itemsize, dtype=np.uint8) if buf is None else buf
            f._read_into(f._abs(addr) + start * dt.itemsize, out)
            return out.view(dt)
        if kind == "implicit":
            if len(self._layout[1]) != 1:
                raise NotImplementedError("implicit chunk index of rank > 1")
            return np.frombuffer(f._mm, dtype=dt, count=n,
                                 offset=f._abs(self._layout[2]) + start * dt.itemsize).copy()
        _, chunk_shape, chunks = self._layout
        if len(chunk_shape) != 1:
            if start != 0 or stop != self.size:
                raise NotImplementedError("row slices of chunked datasets of rank > 1")
            return self._read_nd_chunked(chunk_shape, chunks)
        # read as bytes, chunk by chunk; rows of chunks never written are 0
        isz = dt.itemsize
        out = np.empty(n * isz, dtype=np.uint8) if buf is None else buf
        rows = chunk_shape[0]
        covered = 0
        for off, addr, nbytes in chunks:
            lo, hi = max(off[0], start), min(off[0] + rows, stop, self.size)
            if lo >= hi:
                continue
            if nbytes < rows * isz:
                raise H5FormatError(f"chunk of {nbytes} bytes at {addr} is too small")
            f._read_into(f._abs(addr) + (lo - off[0]) * isz,
                         out[(lo - start) * isz:(hi - start) * isz])
            covered += hi - lo
        if covered < n:
            hit = np.zeros(n, dtype=bool)
            for off, _, _ in chunks:
                lo, hi = max(off[0], start), min(off[0] + rows, stop)
                if lo < hi:
                    hit[lo - start:hi - start] = True
            out.reshape(n, isz)[~hit] = 0
        return out.view(dt)

    def _read_nd_chunked(self, chunk_shape, chunks):
        out = np.zeros(self.shape, dtype=self.dtype)
        for off, addr, nbytes in chunks:
            src = np.frombuffer(self.file._mm, dtype=self.dtype, count=int(np.prod(chunk_shape)),
                                offset=self.file._abs(addr)).reshape(chunk_shape)
            dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, chunk_shape, self.shape))
            out[dst] = src[tuple(slice(0, d.stop - d.start) for d in dst)]
        return out.reshape(-1)

    def read(self, start=0, stop=None, field=None, buf=None):
        """Rows [start, stop) (all by default) as a numpy array in native
        byte order; `field` picks one field of a compound dataset.
        Variable-length strings come back as an object array of bytes.
        buf (optional): a uint8 host array the rows are read into (and that
        the result may be a view of) for 1-D contiguous or chunked data."""
        size = self.size
        stop = size if stop is None else min(int(stop), size)
        start = max(0, min(int(start), stop))
        raw = self._raw(start, stop, buf)
        if self._vlen is not None:
            return self.file._vlen_decode(raw, self._vlen)
        if field is not None:
            raw = raw[field]
        out = raw if raw.dtype.isnative else raw.astype(raw.dtype.newbyteorder("="))
        if not self.shape:
            return out.reshape(())
        if len(self.shape) > 1 and start == 0 and stop == size:
            return out.reshape(self.shape)
        return out


class File:
    """A read-only HDF5 file: ``File(path)[name]`` is a :class:`Dataset` of
    the root group; ``File(path).keys()`` lists the root group's links."""

    def __init__(self, path):
        self.path = os.fspath(path)
        self._fh = open(self.path, "rb")
        try:
            self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError as exc:  # empty file
            self._fh.close()
            raise H5FormatError(f"{self.path} is not an HDF5 file") from exc
        self._gheaps = {}
        try:
            self._superblock()
            self._root = self._group(self._root_addr)
        except Exception:
            self.close()
            raise

    def close(self):
        mm, self._mm = getattr(self, "_mm", None), None
        if mm is not None:
            mm.close()
        self._fh.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def keys(self):
        return list(self._root)

    def __contains__(self, name):
        return name in self._root

    def __getitem__(self, name):
        if name not in self._root:
            raise KeyError(f"no object '{name}' in the root group of {self.path}")
        return Dataset(self, name, self._messages(self._root[name]))

    # -- addresses and superblock -----------------------------------------
    def _abs(self, addr):
        a = self._base + addr
        if a >= len(self._mm):
            raise H5FormatError(f"address {addr} lies beyond the end of the file")
        return a

    def _addr(self, pos):
        return _uint(self._mm, pos, self._so)

    def _read_into(self, pos, buf):
        """Fill the uint8 array buf from absolute file position pos.  Bulk
        data go through read() rather than the memory map, which would fault
        in every page separately."""
        if pos + buf.size > len(self._mm):
            raise H5FormatError(f"data at {pos} run past the end of the file")
        self._fh.seek(pos)
        view = memoryview(buf)
        got = 0
        while got < buf.size:
            k = self._fh.readinto(view[got:])
            if not k:
                raise H5FormatError("unexpected end of file")
            got += k

    def _len(self, pos):
        return _uint(self._mm, pos, self._sl)

    def _superblock(self):
        mm = self._mm
        at = 0
        while at + 8 <= len(mm) and mm[at:at + 8] != SIGNATURE:
            at = 512 if at == 0 else at * 2
        if at + 8 > len(mm):
            raise H5FormatError(f"{self.path} is not an HDF5 file")
        ver = mm[at + 8]
        if ver in (0, 1):
            self._so, self._sl = mm[at + 13], mm[at + 14]
            p = at + (28 if ver == 1 else 24)
            self._base = _uint(mm, p, self._so) or at
            p += 4 * self._so
            # root symbol table entry: link name offset, object header address
            self._root_addr = _uint(mm, p + self._so, self._so)
        elif ver in (2, 3):
            self._so, self._sl = mm[at + 9], mm[at + 10]
            p = at + 12
            self._base = _uint(mm, p, self._so) or at
            self._root_addr = _uint(mm, p + 3 * self._so, self._so)
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")
        if self._so not in (2, 4, 8) or self._sl not in (2, 4, 8):
            raise H5FormatError("bad sizes of offsets/lengths")

    # -- object headers ------------------------------------------------------
    def _messages(self, addr):
        """[(message type, data bytes)] of the object header at addr,
        following continuation messages."""
        mm, pos = self._mm, self._abs(addr)
        out = []
        if mm[pos:pos + 4] == b"OHDR":
            flags = mm[pos + 5]
            p = pos + 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
            w = 1 << (flags & 3)
            size0 = _uint(mm, p, w)
            p += w
            blocks = [(p, p + size0)]
            hdr = 6 if flags & 0x04 else 4
            while blocks:
                p, end = blocks.pop(0)
                while p + hdr <= end:
                    mtype, msize = mm[p], _uint(mm, p + 1, 2)
                    data = mm[p + hdr:p + hdr + msize]
                    p += hdr + msize
                    if mtype == 0x10:
                        c = self._abs(self._addr_of(data, 0))
                        clen = _uint(data, self._so, self._sl)
                        if mm[c:c + 4] != b"OCHK":
                            raise H5FormatError("bad object header continuation block")
                        blocks.append((c + 4, c + clen - 4))
                    elif mtype:
                        out.append((mtype, data))
            return out
        if mm[pos] != 1:
            raise NotImplementedError(f"object header version {mm[pos]}")
        nmsgs = _uint(mm, pos + 2, 2)
        blocks = [(pos + 16, pos + 16 + _uint(mm, pos + 8, 4))]
        seen = 0
        while blocks and seen < nmsgs:
            p, end = blocks.pop(0)
            while p + 8 <= end and seen < nmsgs:
                mtype, msize = _uint(mm, p, 2), _uint(mm, p + 2, 2)
                data = mm[p + 8:p + 8 + msize]
                p += 8 + msize
                seen += 1
                if mtype == 0x10:
                    c = self._abs(self._addr_of(data, 0))
                    blocks.append((c, c + _uint(data, self._so, self._sl)))
                elif mtype:
                    out.append((mtype, data))
        return out

    def _addr_of(self, data, pos):
        return _uint(data, pos, self._so)

    # -- groups -------------------------------------------------------------
    def _group(self, addr):
        links = {}
        for mtype, data in self._messages(addr):
            if mtype == 0x11:  # symbol table: v1 B-tree + local heap
                btree = self._addr_of(data, 0)
                heap = self._addr_of(data, self._so)
                self._walk_symbols(btree, self._heap_data(heap), links)
            elif mtype == 0x06:  # link message
                name, target = self._link(data)
                if target is not None:
                    links[name] = target
            elif mtype == 0x02:  # link info
                flags = data[1]
                p = 2 + (8 if flags & 1 else 0)
                if self._addr_of(data, p) != UNDEF:
                    raise NotImplementedError("groups with dense (fractal heap) link storage")
        return links

    def _heap_data(self, addr):
        mm, pos = self._mm, self._abs(addr)
        if mm[pos:pos + 4] != b"HEAP":
            raise H5FormatError("bad local heap")
        return self._abs(_uint(mm, pos + 8 + 2 * self._sl, self._so))

    def _walk_symbols(self, addr, heap, links):
        mm, pos = self._mm, self._abs(addr)
        if mm[pos:pos + 4] != b"TREE" or mm[pos + 4] != 0:
            raise H5FormatError("bad group B-tree node")
        level, n = mm[pos + 5], _uint(mm, pos + 6, 2)
        p = pos + 8 + 2 * self._so + self._sl  # skip the left sibling, right sibling, key 0
        for _ in range(n):
            child = self._addr(p)
            p += self._so + self._sl
            if level:
                self._walk_symbols(child, heap, links)
                continue
            s = self._abs(child)
            if mm[s:s + 4] != b"SNOD":
                raise H5FormatError("bad symbol table node")
            esize = 2 * self._so + 24
            for k in range(_uint(mm, s + 6, 2)):
                e = s + 8 + k * esize
                name, _ = _cstring(mm, heap + self._addr(e))
                links[name] = self._addr(e + self._so)

    def _link(self, data):
        flags = data[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = data[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        w = 1 << (flags & 3)
        nlen = _uint(data, p, w)
        p += w
        name = bytes(data[p:p + nlen]).decode("utf-8")
        p += nlen
        return name, (self._addr_of(data, p) if ltype == 0 else None)

    # -- dataset messages ----------------------------------------------------
    def _dataspace(self, data):
        ver, rank, flags = data[0], data[1], data[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            if data[3] == 2:  # null dataspace
                return (0,), (0,)
            p = 4
        else:
            raise NotImplementedError(f"dataspace version {ver}")
        dims = tuple(_uint(data, p + k * self._sl, self._sl) for k in range(rank))
        p += rank * self._sl
        maxdims = dims
        if flags & 1:
            maxdims = tuple(None if _uint(data, p + k * self._sl, self._sl) == (1 << 8 * self._sl) - 1
                            else _uint(data, p + k * self._sl, self._sl) for k in range(rank))
        return dims, maxdims

    def _layout(self, data):
        ver, cls = data[0], data[1]
        if ver not in (3, 4):
            raise NotImplementedError(f"data layout message version {ver}")
        if cls == 0:
            size = _uint(data, 2, 2)
            return ("compact", bytes(data[4:4 + size]))
        if cls == 1:
            return ("contiguous", self._addr_of(data, 2))
        if cls != 2:
            raise NotImplementedError(f"data layout class {cls}")
        if ver == 3:
            ndims = data[2]
            btree = self._addr_of(data, 3)
            p = 3 + self._so
            dims = [_uint(data, p + 4 * k, 4) for k in range(ndims)]
            chunks = [] if btree == UNDEF else self._walk_chunks(btree, ndims)
            return ("chunked", tuple(dims[:-1]), chunks)
        flags, ndims, enc = data[2], data[3], data[4]
        p = 5
        dims = [_uint(data, p + enc * k, enc) for k in range(ndims)]
        p += enc * ndims
        index = data[p]
        p += 1
        shape = tuple(dims[:-1])
        nbytes = int(np.prod(shape)) * dims[-1]
        if index == 1:  # single chunk
            if flags & 2:
                raise NotImplementedError("filtered single-chunk datasets")
            addr = self._addr_of(data, p)
            return ("chunked", shape, [] if addr == UNDEF else [((0,) * len(shape), addr, nbytes)])
        if index == 2:  # implicit: every chunk stored in order from one address
            return ("implicit", shape, self._addr_of(data, p))
        names = {3: "fixed array", 4: "extensible array", 5: "version 2 B-tree"}
        raise NotImplementedError(
            f"HDF5 chunk index '{names.get(index, index)}' (files written with libver='latest'); "
            "rewrite the file with h5py's default file format")

    def _walk_chunks(self, addr, ndims):
        """[(chunk offset, address, size in bytes)] of a version 1 B-tree of
        chunks."""
        mm, out, todo = self._mm, [], [addr]
        ksize = 8 + 8 * ndims
        while todo:
            pos = self._abs(todo.pop())
            if mm[pos:pos + 4] != b"TREE" or mm[pos + 4] != 1:
                raise H5FormatError("bad chunk B-tree node")
            level, n = mm[pos + 5], _uint(mm, pos + 6, 2)
            p = pos + 8 + 2 * self._so
            for _ in range(n):
                nbytes, mask = _uint(mm, p, 4), _uint(mm, p + 4, 4)
                off = tuple(_uint(mm, p + 8 + 8 * k, 8) for k in range(ndims - 1))
                child = self._addr(p + ksize)
                if level:
                    todo.append(child)
                else:
                    if mask:
                        raise NotImplementedError("filtered chunks")
                    out.append((off, child, nbytes))
                p += ksize + self._so
        out.sort()
        return out

    # -- variable-length data -------------------------------------------------
    def _global_heap(self, addr):
        objs = self._gheaps.get(addr)
        if objs is not None:
            return objs
        mm, pos = self._mm, self._abs(addr)
        if mm[pos:pos + 4] != b"GCOL":
            raise H5FormatError("bad global heap collection")
        end = pos + self._len(pos + 8)
        p = pos + 8 + self._sl
        objs = {}
        while p + 8 + self._sl <= end:
            idx = _uint(mm, p, 2)
            size = self._len(p + 8)
            if idx == 0:
                break
            objs[idx] = bytes(mm[p + 8 + self._sl:p + 8 + self._sl + size])
            p += 8 + self._sl + ((size + 7) // 8) * 8
        self._gheaps[addr] = objs
        return objs

    def _vlen_decode(self, raw, info):
        kind, base = info
        buf = raw.tobytes()
        isz = raw.dtype.itemsize
        out = np.empty(len(raw), dtype=object)
        for k in range(len(raw)):
            p = k * isz
            n, coll = _uint(buf, p, 4), _uint(buf, p + 4, self._so)
            idx = _uint(buf, p + 4 + self._so, 4)
            data = b"" if n == 0 else self._global_heap(coll)[idx]
            out[k] = data[:n] if kind == "str" else np.frombuffer(data, dtype=base, count=n)
        return out


# ---------------------------------------------------------------------------
# writing
_O = 8  # size of offsets
_L = 8  # size of lengths
_GROUP_LEAF_K = 4
_GROUP_NODE_K = 16
_CHUNK_K = 32


def _pad8(b):
    return b + b"\x00" * (-len(b) % 8)


def _msg(mtype, data, flags=0):
    data = _pad8(data)
    return struct.pack("<HHB3x", mtype, len(data), flags) + data


def _obj_header(msgs):
    body = b"".join(msgs)
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


def _dtype_f64():
    # class 1 (float) version 1, little endian, implied-msb mantissa, sign bit 63
    return struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11,
                                                                      0, 52, 1023)


def _dtype_compound(names):
    head = struct.pack("<BHBI", 0x16, len(names), 0, 8 * len(names))
    body = b""
    for k, name in enumerate(names):
        body += _pad8(name.encode("utf-8") + b"\x00")
        body += struct.pack("<IB3xII16x", 8 * k, 0, 0, 0) + _dtype_f64()
    return head + body


def _dtype_string(size):
    # class 3 version 1, null padded, ASCII
    return struct.pack("<BBBBI", 0x13, 0x01, 0, 0, size)


def _dataspace(n, unlimited):
    if unlimited:
        return struct.pack("<BBBB4xQQ", 1, 1, 1, 0, n, UNDEF)
    return struct.pack("<BBBB4xQ", 1, 1, 0, 0, n)


def _fill_value(alloc_time):
    # version 2, allocation time, fill write time "if set", no fill value defined
    return struct.pack("<BBBB", 2, alloc_time, 2, 0)


def write_table(path, columns, header_lines, name="samples", chunk_rows=None):
    """Write a float64 table as the reference's writer lays it out: the
    compound dataset `name` (one float64 field per column, in order) and
    ``name + '.__table_column_meta__'`` holding `header_lines` (the YAML
    table header) as fixed-length strings.

    columns: list of (column name, 1-D float64 array), all of one length."""
    names = [c for c, _ in columns]
    n = len(columns[0][1]) if columns else 0
    if any(len(v) != n for _, v in columns):
        raise ValueError("columns differ in length")
    if len(set(names)) != len(names) or not names:
        raise ValueError("column names must be unique and non-empty")
    itemsize = 8 * len(names)
    if chunk_rows is None:
        chunk_rows = max(1024, -(-n // (2 * _CHUNK_K)))
    chunk_rows = int(min(max(1, chunk_rows), (1 << 32) // itemsize - 1))
    nchunks = -(-n // chunk_rows) if n else 0
    if nchunks > 2 * _CHUNK_K:
        raise ValueError(f"{nchunks} chunks exceed one B-tree leaf; pass a larger chunk_rows")
    meta_name = name + ".__table_column_meta__"
    enc = [ln.encode("utf-8") for ln in header_lines]
    slen = max([1] + [len(b) for b in enc])
    meta = b"".join(b.ljust(slen, b"\x00") for b in enc)

    # Layout: superblock | root header | heap header | heap data | group B-tree | SNOD
    #         | dataset headers | chunk B-tree | header strings | chunks
    sb_size = 24 + 4 * _O + (2 * _O + 24)
    root_hdr = sb_size
    root_size = 16 + 8 + 2 * _O
    heap_hdr = root_hdr + root_size
    heap_data = heap_hdr + 32
    heap = bytearray(8)  # offset 0: the empty name
    offs = {}
    for nm in sorted([name, meta_name]):
        offs[nm] = len(heap)
        heap += _pad8(nm.encode("utf-8") + b"\x00")
    heap = bytes(heap)
    gtree = heap_data + len(heap)
    gtree_size = 8 + 2 * _O + (2 * _GROUP_NODE_K + 1) * _L + 2 * _GROUP_NODE_K * _O
    snod = gtree + gtree_size
    snod_size = 8 + 2 * _GROUP_LEAF_K * (2 * _O + 24)
    ds_hdr = snod + snod_size

    ctree_size = 8 + 2 * _O + (2 * _CHUNK_K + 1) * (8 + 16) + 2 * _CHUNK_K * _O

    def ds_header(ctree):
        return _obj_header([
            _msg(0x0001, _dataspace(n, True)),
            _msg(0x0003, _dtype_compound(names), flags=1),
            _msg(0x0005, _fill_value(3), flags=1),
            _msg(0x0008, struct.pack("<BBBQII", 3, 2, 2, ctree, chunk_rows, itemsize)),
        ])

    ds_size = len(ds_header(0))
    meta_hdr = ds_hdr + ds_size

    def meta_header(addr):
        return _obj_header([
            _msg(0x0001, _dataspace(len(enc), False)),
            _msg(0x0003, _dtype_string(slen), flags=1),
            _msg(0x0005, _fill_value(2), flags=1),
            _msg(0x0008, struct.pack("<BBQQ", 3, 1, addr, len(meta))),
        ])

    meta_size = len(meta_header(0))
    ctree = meta_hdr + meta_size
    meta_addr = ctree + ctree_size
    chunk0 = meta_addr + len(meta)
    chunk0 += -chunk0 % 8
    chunk_bytes = chunk_rows * itemsize
    eof = chunk0 + nchunks * chunk_bytes

    sb = SIGNATURE + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, _O, _L, 0, _GROUP_LEAF_K,
                                 _GROUP_NODE_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", gtree, heap_hdr)
    root = _obj_header([_msg(0x0011, struct.pack("<QQ", gtree, heap_hdr))])
    heap_head = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), 1, heap_data)
    gnode = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
    gnode += struct.pack("<QQQ", 0, snod, offs[meta_name])
    gnode = gnode.ljust(gtree_size, b"\x00")
    sn = b"SNOD" + struct.pack("<BBH", 1, 0, 2)
    sn += struct.pack("<QQII16x", offs[name], ds_hdr, 0, 0)
    sn += struct.pack("<QQII16x", offs[meta_name], meta_hdr, 0, 0)
    sn = sn.ljust(snod_size, b"\x00")
    cnode = b"TREE" + struct.pack("<BBHQQ", 1, 0, nchunks, UNDEF, UNDEF)
    for k in range(nchunks):
        cnode += struct.pack("<IIQQ", chunk_bytes, 0, k * chunk_rows, 0)
        cnode += struct.pack("<Q", chunk0 + k * chunk_bytes)
    cnode += struct.pack("<IIQQ", 0, 0, nchunks * chunk_rows, itemsize)
    cnode = cnode.ljust(ctree_size, b"\x00")
    head = (sb + root + heap_head + heap + gnode + sn + ds_header(ctree) + meta_header(meta_addr)
            + cnode + meta)
    assert len(head) == chunk0 - (-(meta_addr + len(meta)) % 8) and len(sb) == sb_size
    head = head.ljust(chunk0, b"\x00")

    dt = np.dtype({"names": names, "formats": ["<f8"] * len(names)})
    tmp = os.fspath(path) + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(head)
        for k in range(nchunks):
            lo, hi = k * chunk_rows, min((k + 1) * chunk_rows, n)
            block = np.zeros(chunk_rows, dtype=dt)
            for c, v in columns:
                block[c][:hi - lo] = np.asarray(v[lo:hi], dtype=np.float64)
            fh.write(block.tobytes())
    os.replace(tmp, path)
