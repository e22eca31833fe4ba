"""HDF5 files for run outputs, written without an HDF5 library.

The reference writes `light.h5` (src/output.jl:10-32) and `events.h5`
(scripts/shoot_gammas.jl:12-17,60-78) through HDF5.jl.  This module writes the same
datasets and root attributes in the HDF5 file format itself, so that h5py, HDF5.jl and
the HDF5 tools read the files as they read the reference's.  It uses the oldest layout
every HDF5 library since 1.0 reads, all of it fixed-size with no checksums:

* superblock version 1, 8-byte offsets and lengths;
* the root group as a symbol-table group: a local heap with the link names, one leaf
  node of a version-1 B-tree and one symbol-table node holding every entry, sorted by
  name (the group leaf-node K in the superblock is raised so that one node holds all);
* each dataset as a version-1 object header (dataspace, datatype, fill value, layout)
  with contiguous storage, or chunked storage indexed by one leaf node of a version-1
  B-tree of raw-data chunks (the superblock's indexed-storage K is raised so that one
  node indexes every chunk);
* root attributes as attribute messages in the root group's object header: integers
  as 64-bit (Julia's Int), Float32 and Float64 as such, strings as fixed-length UTF-8
  (HDF5.jl's String datatype).

Arrays are given in the order h5py shows them: a Julia (column-major) array of size
(nx, ny, n) is the C-order (n, ny, nx) array with the same bytes.  The reference chunks
the event datasets along the event axis, up to 4096 events per chunk with the full
other dimensions (scripts/shoot_gammas.jl:11-17); `write(..., chunk_events=4096)` does
the same.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"
GROUP_INTERNAL_K = 16          # the library default; sets the B-tree node size
HEAP_FREE_NULL = 1             # free-list terminator of a local heap
SUPERBLOCK_SIZE = 104          # version 1 (100 bytes), padded to 8
ENTRY_SIZE = 40                # symbol-table entry with 8-byte offsets
ISTORE_K_DEFAULT = 32          # the library default indexed-storage K
ISTORE_K_MAX = 32767           # the largest the library accepts (2K entries per node)

MSG_DATASPACE, MSG_DATATYPE, MSG_FILL, MSG_LAYOUT, MSG_ATTR, MSG_STAB = 0x1, 0x3, 0x5, 0x8, 0xC, 0x11


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _datatype(dt) -> bytes:
    """Datatype message body for a numpy dtype, or ("str", nbytes) for a string."""
    if isinstance(dt, tuple):
        return struct.pack("<B3BI", 0x13, 0x10, 0, 0, dt[1])          # null-terminated UTF-8
    dt = np.dtype(dt)
    if dt.byteorder == ">":
        raise ValueError("big-endian arrays are not written")
    if dt.kind in "iu":
        return struct.pack("<B3BIHH", 0x10, 0x08 if dt.kind == "i" else 0, 0, 0,
                           dt.itemsize, 0, 8 * dt.itemsize)
    if dt.kind == "f" and dt.itemsize in (4, 8):
        sign, eloc, esz, msz, bias = (31, 23, 8, 23, 127) if dt.itemsize == 4 else (63, 52, 11, 52, 1023)
        return struct.pack("<B3BIHHBBBBI", 0x11, 0x20, sign, 0, dt.itemsize, 0,
                           8 * dt.itemsize, eloc, esz, 0, msz, bias)
    raise ValueError(f"no HDF5 datatype for {dt}")


def _dataspace(shape) -> bytes:
    """Version-1 dataspace: rank 0 is a scalar, otherwise fixed dimensions."""
    return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", d) for d in shape)


def _fill_value(itemsize: int, alloc: int = 2) -> bytes:
    """Fill value version 2: allocation late (2) or incremental (3, chunked), written if
    set, defined as zeros."""
    return struct.pack("<BBBBI", 2, alloc, 2, 1, itemsize) + b"\0" * itemsize


def _layout(addr: int, size: int) -> bytes:
    """Layout version 3, contiguous: data address and size."""
    return struct.pack("<BBQQ", 3, 1, addr, size)


def _layout_chunked(btree_addr: int, chunk, itemsize: int) -> bytes:
    """Layout version 3, chunked: the chunk B-tree's address and the chunk dimensions
    followed by the element size."""
    dims = tuple(chunk) + (itemsize,)
    return struct.pack("<BBBQ", 3, 2, len(dims), btree_addr) + b"".join(struct.pack("<I", d) for d in dims)


def _chunk_btree(chunk, shape, chunk_bytes: int, addrs, ik: int) -> bytes:
    """A leaf node of a version-1 B-tree of raw-data chunks (node type 1) indexing chunks
    that split the first dimension only: key i = (chunk bytes, filter mask 0, offset of
    chunk i in every dimension and a trailing 0), child i = its address; the final key
    lies one chunk beyond the last.  Sized for 2*ik entries, as readers expect."""
    rank = len(shape)
    key = lambda nbytes, first: struct.pack("<II", nbytes, 0) + struct.pack(
        "<" + "Q" * (rank + 1), first, *([0] * rank))
    out = b"TREE" + struct.pack("<BBHQQ", 1, 0, len(addrs), UNDEF, UNDEF)
    for i, a in enumerate(addrs):
        out += key(chunk_bytes, i * chunk[0]) + struct.pack("<Q", a)
    out += key(0, len(addrs) * chunk[0])
    return out + b"\0" * (_chunk_btree_size(rank, ik) - len(out))


def _chunk_btree_size(rank: int, ik: int) -> int:
    keysize = 8 + 8 * (rank + 1)
    return 24 + (2 * ik + 1) * keysize + 2 * ik * 8


def _attribute(name: str, value) -> bytes:
    """Attribute message version 1 for a scalar int, float, numpy scalar or string."""
    if isinstance(value, str):
        raw = value.encode("utf-8")
        dt, data = ("str", max(len(raw), 1)), raw if raw else b"\0"
    elif isinstance(value, (bool, np.bool_)):
        raise ValueError(f"attribute {name}: booleans are not written")
    elif isinstance(value, np.generic):
        dt, data = value.dtype, value.tobytes()
    elif isinstance(value, int):
        dt, data = np.dtype("<i8"), struct.pack("<q", value)
    elif isinstance(value, float):
        dt, data = np.dtype("<f8"), struct.pack("<d", value)
    else:
        raise ValueError(f"attribute {name}: unsupported value {value!r}")
    nm = name.encode("utf-8") + b"\0"
    t, s = _datatype(dt), _dataspace(())
    return struct.pack("<BBHHH", 1, 0, len(nm), len(t), len(s)) + _pad8(nm) + _pad8(t) + _pad8(s) + data


def _object_header(messages) -> bytes:
    """Version-1 object header with one chunk: (type, flags, body) per message."""
    chunk = b"".join(struct.pack("<HHB3x", typ, len(_pad8(body)), flags) + _pad8(body)
                     for typ, flags, body in messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(chunk)) + chunk


def write(path, datasets: dict, attrs: dict | None = None, chunk_events: int = 0) -> None:
    """Write `datasets` (name -> array, in h5py order) under the root group and `attrs`
    as root attributes to the HDF5 file `path`.  chunk_events > 0 stores every array of
    rank >= 1 in chunks of that many entries of its first (event) axis with the full other
    dimensions; 0 stores them contiguous."""
    attrs = attrs or {}
    arrays = {}
    for name, a in datasets.items():
        if not name or "/" in name or "\0" in name:
            raise ValueError(f"bad dataset name {name!r}")
        a = np.ascontiguousarray(a)
        _datatype(a.dtype)
        arrays[name] = a
    # chunk shape of each chunked dataset (first axis split, the rest whole); the
    # indexed-storage K lets one leaf node index the most chunks any dataset has
    chunks, most = {}, 0
    if chunk_events > 0:
        for name, a in arrays.items():
            if a.ndim >= 1 and a.size:
                c0 = min(int(chunk_events), a.shape[0])
                chunks[name] = (c0,) + a.shape[1:]
                most = max(most, -(-a.shape[0] // c0))
    ik = max(ISTORE_K_DEFAULT, -(-most // 2))
    if ik > ISTORE_K_MAX:
        raise ValueError(f"{most} chunks exceed the {2 * ISTORE_K_MAX} one B-tree node indexes")
    names = sorted(arrays, key=lambda s: s.encode("utf-8"))
    nsym = len(names)
    leaf_k = max(4, (nsym + 1) // 2)

    # Local heap: "" at offset 0, then the names, each null-terminated to a multiple of 8.
    heap, offs = bytearray(8), {}
    for n in names:
        offs[n] = len(heap)
        heap += _pad8(n.encode("utf-8") + b"\0")

    root_msgs = [(MSG_STAB, 0, struct.pack("<QQ", 0, 0))] + [(MSG_ATTR, 0, _attribute(k, v))
                                                             for k, v in attrs.items()]
    root_oh_size = len(_object_header(root_msgs))
    btree_size = 24 + 2 * GROUP_INTERNAL_K * 8 + (2 * GROUP_INTERNAL_K + 1) * 8
    snod_size = 8 + 2 * leaf_k * ENTRY_SIZE

    root_addr = SUPERBLOCK_SIZE
    heap_addr = root_addr + root_oh_size
    heap_data_addr = heap_addr + 32
    btree_addr = heap_data_addr + len(heap)
    snod_addr = btree_addr + btree_size
    pos = snod_addr + snod_size

    headers, data_addr, btree_addr_of, chunk_addrs = {}, {}, {}, {}
    for n in names:
        a = arrays[n]
        if n in chunks:
            lay = _layout_chunked(0, chunks[n], a.dtype.itemsize)
            fill = _fill_value(a.dtype.itemsize, alloc=3)
        else:
            lay, fill = _layout(0, 0), _fill_value(a.dtype.itemsize)
        msgs = [(MSG_DATASPACE, 0, _dataspace(a.shape)), (MSG_DATATYPE, 1, _datatype(a.dtype)),
                (MSG_FILL, 1, fill), (MSG_LAYOUT, 0, lay)]
        headers[n] = (pos, msgs)
        pos += len(_object_header(msgs))
    for n in names:
        a = arrays[n]
        if n in chunks:
            btree_addr_of[n] = pos
            pos += _chunk_btree_size(a.ndim, ik)
            pos += -pos % 8
            cb = int(np.prod(chunks[n])) * a.dtype.itemsize
            nch = -(-a.shape[0] // chunks[n][0])
            chunk_addrs[n] = [pos + i * (cb + (-cb % 8)) for i in range(nch)]
            pos += nch * (cb + (-cb % 8))
        else:
            data_addr[n] = pos if a.nbytes else UNDEF
            pos += a.nbytes + (-a.nbytes % 8)
    eof = pos

    root_msgs[0] = (MSG_STAB, 0, struct.pack("<QQ", btree_addr, heap_addr))
    sb = (SIGNATURE + struct.pack("<BBBBBBBBHHI", 1, 0, 0, 0, 0, 8, 8, 0, leaf_k, GROUP_INTERNAL_K, 0)
          + struct.pack("<HH", ik, 0)
          + struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
          + struct.pack("<QQII", 0, root_addr, 1, 0) + struct.pack("<QQ", btree_addr, heap_addr))
    sb += b"\0" * (SUPERBLOCK_SIZE - len(sb))
    assert len(sb) == SUPERBLOCK_SIZE
    heap_hdr = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), HEAP_FREE_NULL, heap_data_addr)

    # One leaf B-tree node: key 0 = "" (heap offset 0), child 0 = the symbol-table node,
    # key 1 = the largest name.  Unused key and child slots stay zero.
    btree = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1 if nsym else 0, UNDEF, UNDEF)
    if nsym:
        btree += struct.pack("<QQQ", 0, snod_addr, offs[names[-1]])
    btree += b"\0" * (btree_size - len(btree))
    snod = b"SNOD" + struct.pack("<BBH", 1, 0, nsym)
    for n in names:
        snod += struct.pack("<QQII16x", offs[n], headers[n][0], 0, 0)
    snod += b"\0" * (snod_size - len(snod))

    with open(Path(path), "wb") as f:
        f.write(sb)
        f.write(_object_header(root_msgs))
        f.write(heap_hdr)
        f.write(bytes(heap))
        f.write(btree)
        f.write(snod)
        for n in names:
            a = arrays[n]
            addr, msgs = headers[n]
            assert f.tell() == addr
            if n in chunks:
                msgs[-1] = (MSG_LAYOUT, 0, _layout_chunked(btree_addr_of[n], chunks[n], a.dtype.itemsize))
            else:
                msgs[-1] = (MSG_LAYOUT, 0, _layout(data_addr[n], a.nbytes))
            f.write(_object_header(msgs))
        for n in names:
            a = arrays[n]
            if n in chunks:
                assert f.tell() == btree_addr_of[n]
                c0 = chunks[n][0]
                cb = int(np.prod(chunks[n])) * a.dtype.itemsize
                f.write(_chunk_btree(chunks[n], a.shape, cb, chunk_addrs[n], ik))
                f.write(b"\0" * (-f.tell() % 8))
                flat = a.reshape(a.shape[0], -1)
                for i, ca in enumerate(chunk_addrs[n]):
                    assert f.tell() == ca
                    part = flat[i * c0:(i + 1) * c0].tobytes()
                    f.write(part + b"\0" * (cb - len(part)))      # the edge chunk is full-size
                    f.write(b"\0" * (-cb % 8))
            elif a.nbytes:
                assert f.tell() == data_addr[n]
                f.write(a.tobytes())
                f.write(b"\0" * (-a.nbytes % 8))
        assert f.tell() == eof
