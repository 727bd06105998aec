"""HDF5 snapshots in the layout of the reference's ``H5Saver``.

The reference saves the conserved state with h5py
(``simulation/set_simulator.py:126-318``): one group ``save_step{k}`` per
snapshot with the attributes ``time`` (float64) and ``step`` (int64) and a
gzip-compressed dataset ``u`` (12, nx, ny).  h5py is not part of this
framework's dependencies, so this module writes and reads that file format
itself (HDF5 file format specification, version 0 superblock, version 1
object headers, symbol-table groups, chunked datasets indexed by a version 1
B-tree, deflate filter) with numpy and zlib.  Files written here open in
h5py/libhdf5, and files h5py writes with its default settings read here.

Datasets are chunked (~1 MiB chunks of whole rows) and stored through the
HDF5 shuffle and deflate filters (level 4, h5py's gzip default; the reference
uses deflate alone, which leaves floating-point data nearly incompressible;
``shuffle=False`` gives its exact filter set).  Compression runs on a pool
of host threads.

``H5Saver`` mirrors the reference's class (``save``, ``list_snapshots``,
``load``, ``load_all``, ``close``) and adds ``save_async`` (device-to-host
copy on a side stream into pinned memory, compression and writing on a
worker thread, so stepping continues meanwhile) and ``load_state`` (the
state and its (gamma, T) for a restart, which the reference lacks).
"""
import os
import struct
import threading
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"
_HEAP_FREE_NULL = 1          # local heap "no free block" marker
_GROUP_LEAF_K = 4            # symbol-table node: up to 2K entries
_GROUP_INTERNAL_K = 16       # group B-tree node: up to 2K children
_CHUNK_K = 32                # chunk B-tree node: up to 2K children (v0 default)

# object header message types
_MSG_NIL, _MSG_SPACE, _MSG_LINFO, _MSG_DTYPE, _MSG_FILL_OLD, _MSG_FILL = 0, 1, 2, 3, 4, 5
_MSG_LINK, _MSG_LAYOUT, _MSG_FILTER, _MSG_ATTR, _MSG_CONT, _MSG_STAB = 6, 8, 11, 12, 16, 17


def _pad8(b):
    return b + bytes((-len(b)) % 8)


# ---------------------------------------------------------------- encoding
def _dtype_msg(dt):
    dt = np.dtype(dt)
    if dt.byteorder == ">":
        raise ValueError("big-endian data is not supported")
    n = dt.itemsize
    if dt.kind == "f":
        if n == 8:
            sign, eloc, esize, msize, bias = 63, 52, 11, 52, 1023
        elif n == 4:
            sign, eloc, esize, msize, bias = 31, 23, 8, 23, 127
        else:
            raise ValueError("unsupported float size %d" % n)
        # class 1 version 1; little endian, implied mantissa msb
        return (struct.pack("<BBBBI", 0x11, 0x20, sign, 0, n) +
                struct.pack("<HHBBBBI", 0, 8 * n, eloc, esize, 0, msize, bias))
    if dt.kind in "iu":
        return (struct.pack("<BBBBI", 0x10, 0x08 if dt.kind == "i" else 0, 0, 0, n) +
                struct.pack("<HH", 0, 8 * n))
    raise ValueError("unsupported dtype %s" % dt)


def _space_msg(shape):
    return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(
        struct.pack("<Q", int(d)) for d in shape)


def _attr_msg(name, value):
    value = np.asarray(value)
    if value.dtype.kind not in "fiu":
        raise ValueError("attribute %r: numeric values only" % name)
    value = value.astype(value.dtype.newbyteorder("<"))
    nm = name.encode() + b"\0"
    dt = _dtype_msg(value.dtype)
    sp = _space_msg(value.shape)
    return (struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(sp)) + _pad8(nm) + _pad8(dt) +
            _pad8(sp) + value.tobytes())


def _object_header(msgs):
    """Version 1 object header; msgs = [(type, data, flags)]."""
    body = b""
    for t, data, flags in msgs:
        d = _pad8(data)
        body += struct.pack("<HHB3x", t, len(d), flags) + d
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


def _ste(name_off, ohdr, scratch=None):
    """Symbol-table entry; scratch = (btree, heap) for a group."""
    if scratch is None:
        return struct.pack("<QQII16x", name_off, ohdr, 0, 0)
    return struct.pack("<QQIIQQ", name_off, ohdr, 1, 0, scratch[0], scratch[1])


class _Writer:
    """Append-only file with random-access patching."""

    def __init__(self, fh):
        self.fh = fh
        fh.seek(0, os.SEEK_END)
        self.eof = fh.tell()

    def alloc(self, data):
        addr = self.eof
        self.fh.seek(addr)
        self.fh.write(data)
        self.eof += len(data)
        pad = (-self.eof) % 8
        if pad:
            self.fh.write(bytes(pad))
            self.eof += pad
        return addr

    def patch(self, addr, data):
        self.fh.seek(addr)
        self.fh.write(data)


def _write_group(w, children, attrs=None):
    """Group with symbol table; children = [(name, ste_tail)] where ste_tail
    is (ohdr, scratch or None).  Returns (ohdr_addr, btree_addr, heap_addr)."""
    children = sorted(children, key=lambda c: c[0].encode())
    # local heap: "" at offset 0, then the names
    seg = bytearray(8)
    offs = []
    for name, _ in children:
        offs.append(len(seg))
        seg += _pad8(name.encode() + b"\0")
    if len(seg) < 88:                  # libhdf5's minimum heap data size
        seg += bytes(88 - len(seg))
    seg_addr = w.alloc(bytes(seg))
    heap_addr = w.alloc(b"HEAP" + struct.pack("<B3xQQQ", 0, len(seg), _HEAP_FREE_NULL, seg_addr))
    # symbol-table nodes of up to 2*leaf_K entries
    per = 2 * _GROUP_LEAF_K
    snods = []
    for i in range(0, len(children), per):
        part = list(zip(offs[i:i + per], children[i:i + per]))
        body = b"".join(_ste(off, c[1][0], c[1][1]) for off, c in part)
        body += bytes(per * 40 - len(body))
        addr = w.alloc(b"SNOD" + struct.pack("<BBH", 1, 0, len(part)) + body)
        last = part[-1][0] if part else 0
        snods.append((addr, last))
    cap = 2 * _GROUP_INTERNAL_K
    if snods:
        btree = _write_btree(w, 0, snods, cap, key=lambda last: struct.pack("<Q", last),
                             key0=struct.pack("<Q", 0))
    else:                                # empty group: a B-tree node with no children
        blob = b"TREE" + struct.pack("<BBHQQ", 0, 0, 0, UNDEF, UNDEF) + struct.pack("<Q", 0)
        btree = w.alloc(blob + bytes(24 + (cap + 1) * 8 + cap * 8 - len(blob)))
    msgs = [(_MSG_STAB, struct.pack("<QQ", btree, heap_addr), 0)]
    for name, value in (attrs or {}).items():
        msgs.append((_MSG_ATTR, _attr_msg(name, value), 0))
    ohdr = w.alloc(_object_header(msgs))
    return ohdr, btree, heap_addr


def _write_btree(w, ntype, leaves, cap, key, key0, level=0):
    """Version 1 B-tree over `leaves` = [(child_addr, right_key_value)];
    keys: key0 before the first child, key(right) after each child."""
    ksize = len(key0)
    size = 24 + (cap + 1) * ksize + cap * 8
    nodes = []
    for i in range(0, len(leaves), cap):
        part = leaves[i:i + cap]
        k0 = key0 if i == 0 else key(leaves[i - 1][1])
        body = k0 + b"".join(struct.pack("<Q", a) + key(r) for a, r in part)
        hdr = b"TREE" + struct.pack("<BBHQQ", ntype, level, len(part), UNDEF, UNDEF)
        blob = hdr + body
        addr = w.alloc(blob + bytes(size - len(blob)))
        nodes.append((addr, part[-1][1]))
    if len(nodes) == 1:
        return nodes[0][0]
    return _write_btree(w, ntype, nodes, cap, key, key0, level + 1)


def _chunk_shape(shape, itemsize, target=1 << 20):
    """Chunks of whole trailing rows, about `target` bytes each."""
    shape = tuple(int(s) for s in shape)
    if not shape:
        return ()
    ch = [1] * len(shape)
    ch[-1] = max(shape[-1], 1)
    rows = max(1, target // max(ch[-1] * itemsize, 1))
    if len(shape) >= 2:
        ch[-2] = int(min(max(shape[-2], 1), rows))
    return tuple(ch)


def _compress_chunks(arr, chunks, level, pool, shuffle):
    """[(chunk offset tuple, filtered bytes)] in C order of the chunk grid:
    the HDF5 shuffle filter (bytes grouped by significance, which lets
    deflate compress floating-point data) when asked, then deflate."""
    shape = arr.shape
    grid = [range(0, max(s, 1), c) for s, c in zip(shape, chunks)]
    offsets = [()]
    for g in grid:
        offsets = [o + (x,) for o in offsets for x in g]

    def one(off):
        sl = tuple(slice(o, o + c) for o, c in zip(off, chunks))
        block = arr[sl]
        if block.shape != tuple(chunks):          # edge chunk: stored whole
            full = np.zeros(chunks, dtype=arr.dtype)
            full[tuple(slice(0, b) for b in block.shape)] = block
            block = full
        raw = np.ascontiguousarray(block)
        if shuffle and raw.dtype.itemsize > 1:
            raw = raw.view(np.uint8).reshape(-1, raw.dtype.itemsize).T
        return off, zlib.compress(np.ascontiguousarray(raw).tobytes(), level)

    if pool is None:
        return [one(o) for o in offsets]
    return list(pool.map(one, offsets))


def _write_dataset(w, arr, level, pool, chunk_bytes=1 << 20, shuffle=True):
    arr = np.ascontiguousarray(arr)
    arr = arr.astype(arr.dtype.newbyteorder("<"), copy=False)
    if arr.ndim == 0:
        raise ValueError("scalar datasets are not supported")
    chunks = _chunk_shape(arr.shape, arr.dtype.itemsize, chunk_bytes)
    rank = arr.ndim
    parts = _compress_chunks(arr, chunks, level, pool, shuffle)
    leaves = []
    for off, data in parts:
        addr = w.alloc(data)
        leaves.append((addr, (off, len(data))))

    def key(info):
        off, nbytes = info
        return struct.pack("<II", nbytes, 0) + b"".join(struct.pack("<Q", o) for o in off) + \
            struct.pack("<Q", 0)

    # key before child i = that chunk's offset; after the last = its upper bound
    def right_key(i):
        if i + 1 < len(parts):
            return key((parts[i + 1][0], len(parts[i + 1][1])))
        off = parts[i][0]
        return struct.pack("<II", 0, 0) + b"".join(
            struct.pack("<Q", o + c) for o, c in zip(off, chunks)) + struct.pack("<Q", 0)

    ksize = 8 + 8 * (rank + 1)
    cap = 2 * _CHUNK_K
    nodes = []
    size = 24 + (cap + 1) * ksize + cap * 8
    for i in range(0, len(leaves), cap):
        part = list(range(i, min(i + cap, len(leaves))))
        body = key((parts[i][0], len(parts[i][1])))
        for j in part:
            body += struct.pack("<Q", leaves[j][0]) + right_key(j)
        blob = b"TREE" + struct.pack("<BBHQQ", 1, 0, len(part), UNDEF, UNDEF) + body
        nodes.append((w.alloc(blob + bytes(size - len(blob))), part[0], part[-1]))
    level = 0
    while len(nodes) > 1:
        level += 1
        up = []
        for i in range(0, len(nodes), cap):
            part = nodes[i:i + cap]
            body = key((parts[part[0][1]][0], len(parts[part[0][1]][1])))
            for addr, first, last in part:
                body += struct.pack("<Q", addr) + right_key(last)
            blob = b"TREE" + struct.pack("<BBHQQ", 1, level, len(part), UNDEF, UNDEF) + body
            up.append((w.alloc(blob + bytes(size - len(blob))), part[0][1], part[-1][2]))
        nodes = up
    btree = nodes[0][0]
    layout = struct.pack("<BBBQ", 3, 2, rank + 1, btree) + b"".join(
        struct.pack("<I", c) for c in chunks) + struct.pack("<I", arr.dtype.itemsize)
    # filter pipeline (version 1): [shuffle (id 2, element size)], deflate
    # (id 1, level); both marked optional as libhdf5 marks them
    filt = b""
    if shuffle:
        filt += struct.pack("<HHHH", 2, 0, 1, 1) + struct.pack("<I4x", arr.dtype.itemsize)
    filt += struct.pack("<HHHH", 1, 0, 1, 1) + struct.pack("<I4x", level)
    filt = struct.pack("<BB6x", 1, 2 if shuffle else 1) + filt
    msgs = [(_MSG_SPACE, _space_msg(arr.shape), 0),
            (_MSG_DTYPE, _dtype_msg(arr.dtype), 1),
            (_MSG_FILL, struct.pack("<BBBB", 2, 3, 2, 0), 1),
            (_MSG_LAYOUT, layout, 0),
            (_MSG_FILTER, filt, 0)]
    return w.alloc(_object_header(msgs))


def _superblock(root_ohdr, root_btree, root_heap, eof):
    return (SIGNATURE + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0) +
            struct.pack("<HHI", _GROUP_LEAF_K, _GROUP_INTERNAL_K, 0) +
            struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF) +
            _ste(0, root_ohdr, (root_btree, root_heap)))


# ---------------------------------------------------------------- decoding
class _Reader:
    """Reads the subset of HDF5 that h5py's default settings produce."""

    def __init__(self, fh):
        self.fh = fh
        sig = self.read(0, 8)
        if sig != SIGNATURE:
            raise ValueError("not an HDF5 file (no signature at offset 0)")
        ver = self.read(8, 1)[0]
        self.sb_version = ver
        self.base = 0
        if ver in (0, 1):
            so, sl = self.read(13, 2)
            if (so, sl) != (8, 8):
                raise NotImplementedError("only 8-byte offsets and lengths are supported")
            p = 24 if ver == 0 else 28
            self.base, _, self.eof, _ = struct.unpack("<QQQQ", self.read(p, 32))
            self.root_ste_at = p + 32
            _, self.root, ctype, _ = struct.unpack("<QQII", self.read(p + 32, 24))
        elif ver in (2, 3):
            so, sl = self.read(9, 2)
            if (so, sl) != (8, 8):
                raise NotImplementedError("only 8-byte offsets and lengths are supported")
            self.base, _, self.eof, self.root = struct.unpack("<QQQQ", self.read(12, 32))
            self.root_ste_at = None
        else:
            raise NotImplementedError("superblock version %d" % ver)

    def read(self, addr, n):
        self.fh.seek(addr)
        b = self.fh.read(n)
        if len(b) != n:
            raise ValueError("truncated HDF5 file")
        return b

    # ---- object headers
    def messages(self, addr):
        first = self.read(addr, 4)
        if first == b"OHDR":
            return self._messages_v2(addr)
        ver, _, nmsg, _, size = struct.unpack("<BBHII", self.read(addr, 12))
        if ver != 1:
            raise NotImplementedError("object header version %d" % ver)
        out = []
        blocks = [(addr + 16, size)]
        while blocks:
            start, length = blocks.pop(0)
            data = self.read(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsg:
                t, sz, flags = struct.unpack("<HHB", data[p:p + 5])
                body = data[p + 8:p + 8 + sz]
                if flags & 2:
                    raise NotImplementedError("shared object header messages")
                if t == _MSG_CONT:
                    blocks.append(struct.unpack("<QQ", body[:16]))
                out.append((t, body))
                p += 8 + sz
        return out

    def _messages_v2(self, addr):
        hdr = self.read(addr, 6)
        ver, flags = hdr[4], hdr[5]
        if ver != 2:
            raise NotImplementedError("object header version %d" % ver)
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        nsz = 1 << (flags & 3)
        size = int.from_bytes(self.read(p, nsz), "little")
        p += nsz
        blocks = [(p, size)]
        out = []
        track = bool(flags & 0x04)
        while blocks:
            start, length = blocks.pop(0)
            data = self.read(start, length)
            q = 0
            hl = 4 + (2 if track else 0)
            while q + hl <= length:
                t = data[q]
                sz = struct.unpack("<H", data[q + 1:q + 3])[0]
                mflags = data[q + 3]
                body = data[q + hl:q + hl + sz]
                if mflags & 2:
                    raise NotImplementedError("shared object header messages")
                if t == _MSG_CONT:
                    caddr, clen = struct.unpack("<QQ", body[:16])
                    blocks.append((caddr + 4, clen - 8))   # skip "OCHK", checksum
                out.append((t, body))
                q += hl + sz
        return out

    # ---- groups
    def group_entries(self, addr):
        """{name: object header address} of a group."""
        out = {}
        for t, body in self.messages(addr):
            if t == _MSG_STAB:
                btree, heap = struct.unpack("<QQ", body[:16])
                self._stab_entries(btree, heap, out)
            elif t == _MSG_LINK:
                name, target = self._link(body)
                if target is not None:
                    out[name] = target
            elif t == _MSG_LINFO:
                fheap = struct.unpack("<Q", body[2 + (8 if body[1] & 1 else 0):][:8])[0]
                if fheap != UNDEF:
                    raise NotImplementedError("groups with dense link storage")
        return out

    def _stab_entries(self, btree, heap, out):
        hh = self.read(heap, 32)
        if hh[:4] != b"HEAP":
            raise ValueError("bad local heap")
        seg_size, _, seg_addr = struct.unpack("<QQQ", hh[8:32])
        seg = self.read(seg_addr, seg_size)

        def name_at(off):
            return seg[off:seg.index(b"\0", off)].decode()

        def walk(node):
            h = self.read(node, 24)
            if h[:4] != b"TREE" or h[4] != 0:
                raise ValueError("bad group B-tree node")
            level, n = h[5], struct.unpack("<H", h[6:8])[0]
            body = self.read(node + 24, 8 + n * 16)
            for i in range(n):
                child = struct.unpack("<Q", body[8 + 16 * i:16 + 16 * i])[0]
                if level > 0:
                    walk(child)
                    continue
                sn = self.read(child, 8)
                if sn[:4] != b"SNOD":
                    raise ValueError("bad symbol table node")
                cnt = struct.unpack("<H", sn[6:8])[0]
                ents = self.read(child + 8, 40 * cnt)
                for j in range(cnt):
                    off, oh = struct.unpack("<QQ", ents[40 * j:40 * j + 16])
                    out[name_at(off)] = oh
        walk(btree)

    @staticmethod
    def _link(body):
        ver, flags = body[0], body[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = body[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        nlen_size = 1 << (flags & 3)
        nlen = int.from_bytes(body[p:p + nlen_size], "little")
        p += nlen_size
        name = body[p:p + nlen].decode()
        p += nlen
        if ltype != 0:
            return name, None           # soft / external links are not followed
        return name, struct.unpack("<Q", body[p:p + 8])[0]

    # ---- datasets and attributes
    @staticmethod
    def _dtype(body):
        cls, ver = body[0] & 0x0F, body[0] >> 4
        bits = body[1]
        size = struct.unpack("<I", body[4:8])[0]
        if bits & 1:
            raise NotImplementedError("big-endian data")
        if cls == 1:
            return np.dtype("<f%d" % size)
        if cls == 0:
            return np.dtype(("<i%d" if bits & 0x08 else "<u%d") % size)
        raise NotImplementedError("datatype class %d" % cls)

    @staticmethod
    def _space(body):
        ver, rank = body[0], body[1]
        if ver == 1:
            p = 8
            if rank == 0:
                return ()
        elif ver == 2:
            if body[3] in (0, 2):        # scalar / null
                return ()
            p = 4
        else:
            raise NotImplementedError("dataspace version %d" % ver)
        return tuple(struct.unpack("<Q", body[p + 8 * i:p + 8 * i + 8])[0] for i in range(rank))

    def attributes(self, addr):
        out = {}
        for t, body in self.messages(addr):
            if t != _MSG_ATTR:
                continue
            ver = body[0]
            if ver == 1:
                nlen, dlen, slen = struct.unpack("<HHH", body[2:8])
                p = 8
                name = body[p:p + nlen - 1].decode()
                p += (nlen + 7) & ~7
                dt = body[p:p + dlen]
                p += (dlen + 7) & ~7
                sp = body[p:p + slen]
                p += (slen + 7) & ~7
            elif ver in (2, 3):
                nlen, dlen, slen = struct.unpack("<HHH", body[2:8])
                p = 8 + (1 if ver == 3 else 0)
                name = body[p:p + nlen - 1].decode()
                p += nlen
                dt = body[p:p + dlen]
                p += dlen
                sp = body[p:p + slen]
                p += slen
            else:
                raise NotImplementedError("attribute message version %d" % ver)
            dtype = self._dtype(dt)
            shape = self._space(sp)
            n = int(np.prod(shape)) if shape else 1
            val = np.frombuffer(body[p:p + n * dtype.itemsize], dtype=dtype)
            out[name] = val.reshape(shape)[()] if shape else val[0]
        return out

    def dataset(self, addr):
        shape = dtype = layout = None
        filters = []
        for t, body in self.messages(addr):
            if t == _MSG_SPACE:
                shape = self._space(body)
            elif t == _MSG_DTYPE:
                dtype = self._dtype(body)
            elif t == _MSG_LAYOUT:
                layout = body
            elif t == _MSG_FILTER:
                filters = self._filters(body)
        if shape is None or dtype is None or layout is None:
            raise ValueError("object is not a dataset")
        ver, cls = layout[0], layout[1]
        if ver != 3:
            raise NotImplementedError("data layout message version %d" % ver)
        n = int(np.prod(shape)) if shape else 1
        if cls == 0:                     # compact
            size = struct.unpack("<H", layout[2:4])[0]
            return np.frombuffer(layout[4:4 + size], dtype=dtype)[:n].reshape(shape).copy()
        if cls == 1:                     # contiguous
            a, size = struct.unpack("<QQ", layout[2:18])
            if a == UNDEF:
                return np.zeros(shape, dtype)
            return np.frombuffer(self.read(a, n * dtype.itemsize), dtype=dtype).reshape(shape)
        if cls != 2:
            raise NotImplementedError("layout class %d" % cls)
        nd = layout[2]
        btree = struct.unpack("<Q", layout[3:11])[0]
        cdims = struct.unpack("<%dI" % nd, layout[11:11 + 4 * nd])[:-1]
        out = np.zeros(shape, dtype)
        if btree == UNDEF:
            return out
        self._read_chunks(btree, nd, cdims, dtype, filters, out)
        return out

    @staticmethod
    def _filters(body):
        ver, nf = body[0], body[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(nf):
            fid = struct.unpack("<H", body[p:p + 2])[0]
            if ver == 1 or fid >= 256:
                nlen, flags, nvals = struct.unpack("<HHH", body[p + 2:p + 8])
                p += 8
                p += ((nlen + 7) & ~7) if ver == 1 else nlen
            else:
                flags, nvals = struct.unpack("<HH", body[p + 2:p + 6])
                p += 6
            vals = struct.unpack("<%dI" % nvals, body[p:p + 4 * nvals])
            p += 4 * nvals
            if ver == 1 and nvals % 2:
                p += 4
            out.append((fid, vals))
        return out

    def _read_chunks(self, node, nd, cdims, dtype, filters, out):
        h = self.read(node, 24)
        if h[:4] != b"TREE" or h[4] != 1:
            raise ValueError("bad chunk B-tree node")
        level, n = h[5], struct.unpack("<H", h[6:8])[0]
        ks = 8 + 8 * nd
        body = self.read(node + 24, (n + 1) * ks + n * 8)
        for i in range(n):
            k = body[i * (ks + 8):i * (ks + 8) + ks]
            child = struct.unpack("<Q", body[i * (ks + 8) + ks:(i + 1) * (ks + 8)])[0]
            if level > 0:
                self._read_chunks(child, nd, cdims, dtype, filters, out)
                continue
            nbytes, mask = struct.unpack("<II", k[:8])
            off = struct.unpack("<%dQ" % nd, k[8:8 + 8 * nd])[:-1]
            raw = self.read(child, nbytes)
            for j, (fid, vals) in reversed(list(enumerate(filters))):
                if mask & (1 << j):
                    continue
                if fid == 1:
                    raw = zlib.decompress(raw)
                elif fid == 2:
                    sz = vals[0] if vals else dtype.itemsize
                    raw = np.frombuffer(raw, np.uint8).reshape(sz, -1).T.tobytes()
                elif fid == 3:
                    raw = raw[:-4]
                else:
                    raise NotImplementedError("HDF5 filter %d" % fid)
            block = np.frombuffer(raw, dtype=dtype).reshape(cdims)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, cdims, out.shape))
            out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]


# ---------------------------------------------------------------- file
class H5File:
    """An HDF5 file of top-level groups holding numeric datasets and
    numeric scalar attributes (the structure H5Saver writes)."""

    def __init__(self, path, mode="a", gzip_level=4, threads=16, chunk_bytes=1 << 20,
                 shuffle=True):
        self.path = str(path)
        self.level = gzip_level
        self.chunk_bytes = chunk_bytes
        self.shuffle = shuffle
        self._pool = ThreadPoolExecutor(max(1, threads)) if threads > 1 else None
        exists = os.path.exists(self.path) and os.path.getsize(self.path) > 0
        if mode == "r":
            self.fh = open(self.path, "rb")
            self.writable = False
        elif mode in ("a", "r+"):
            if not exists and mode == "r+":
                raise FileNotFoundError(self.path)
            self.fh = open(self.path, "r+b" if exists else "w+b")
            self.writable = True
        elif mode == "w":
            self.fh = open(self.path, "w+b")
            self.writable = True
            exists = False
        else:
            raise ValueError("mode must be 'r', 'r+', 'a' or 'w'")
        self._lock = threading.Lock()
        self.top = {}                     # name -> (ohdr, scratch or None)
        if exists:
            r = _Reader(self.fh)
            if self.writable and r.sb_version != 0:
                raise NotImplementedError("appending to a version %d superblock file"
                                          % r.sb_version)
            for name, oh in r.group_entries(r.root).items():
                self.top[name] = (oh, self._scratch(r, oh))
        elif self.writable:
            self.fh.write(bytes(96))       # superblock, written at flush
            self.flush()

    @staticmethod
    def _scratch(r, oh):
        for t, body in r.messages(oh):
            if t == _MSG_STAB:
                return struct.unpack("<QQ", body[:16])
        return None

    def keys(self):
        return list(self.top)

    def __contains__(self, key):
        return key in self.top

    def write_group(self, name, datasets, attrs):
        """Create (or replace) group `name` with the given datasets and
        attributes; takes effect in the file at the next flush."""
        if not self.writable:
            raise OSError("file is read-only")
        with self._lock:
            w = _Writer(self.fh)
            kids = []
            for dname, arr in datasets.items():
                kids.append((dname, (_write_dataset(w, np.asarray(arr), self.level,
                                                    self._pool, self.chunk_bytes,
                                                    self.shuffle), None)))
            oh, bt, hp = _write_group(w, kids, attrs)
            self.top[name] = (oh, (bt, hp))

    def delete(self, name):
        with self._lock:
            self.top.pop(name, None)

    def flush(self):
        if not self.writable:
            return
        with self._lock:
            w = _Writer(self.fh)
            oh, bt, hp = _write_group(w, list(self.top.items()))
            w.patch(0, _superblock(oh, bt, hp, w.eof))
            self.fh.flush()
            os.fsync(self.fh.fileno())

    def read_group(self, name):
        """(datasets, attributes) of top-level group `name`."""
        with self._lock:
            r = _Reader(self.fh)
            oh = self.top[name][0]
            data = {k: r.dataset(a) for k, a in r.group_entries(oh).items()}
            return data, r.attributes(oh)

    def close(self):
        if self.fh is None:
            return
        if self.writable:
            self.flush()
        self.fh.close()
        self.fh = None
        if self._pool is not None:
            self._pool.shutdown()


# ---------------------------------------------------------------- H5Saver
class H5Saver:
    """``H5Saver`` of simulation/set_simulator.py:126-318 (2-D form)."""

    def __init__(self, filepath, data_dim="2D", simulator=None, gzip_level=4, shuffle=True):
        if data_dim != "2D":
            raise NotImplementedError("janc_amd implements the 2-D solver only")
        self.filepath = str(filepath)
        self.sim = simulator
        self.level = gzip_level
        self.file = H5File(self.filepath, "a", gzip_level, shuffle=shuffle)
        self._worker = ThreadPoolExecutor(1)
        self._pending = []

    def save(self, save_step, t, step, **arrays):
        """set_simulator.py:155-175: group save_step{k}, attributes time and
        step, one gzip dataset per array (device arrays are copied to the
        host first)."""
        self.wait()
        key = "save_step%d" % save_step
        host = {k: _to_host(v) for k, v in arrays.items()}
        self.file.write_group(key, host, {"time": np.float64(float(t)),
                                          "step": np.int64(int(step))})
        self.file.flush()

    def save_async(self, save_step, t, step, **arrays):
        """As save, but the device-to-host copies go on a side stream into
        pinned buffers and the compression and writing run on a worker
        thread; returns at once.  The arrays may be overwritten by later
        device work on the current stream (the copy is ordered before it)."""
        import torch
        key = "save_step%d" % save_step
        staged = {}
        ev = None
        for k, v in arrays.items():
            if isinstance(v, torch.Tensor) and v.is_cuda:
                if ev is None:
                    side = torch.cuda.Stream(device=v.device)
                    side.wait_stream(torch.cuda.current_stream(v.device))
                buf = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
                with torch.cuda.stream(side):
                    buf.copy_(v, non_blocking=True)
                    v.record_stream(side)
                ev = torch.cuda.Event()
                ev.record(side)
                staged[k] = buf
            else:
                staged[k] = _to_host(v)
        if ev is not None:
            # later work on the current stream must not overwrite the source
            # before the copy has read it
            torch.cuda.current_stream().wait_event(ev)
        t = float(t)

        def job():
            if ev is not None:
                ev.synchronize()
            host = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in staged.items()}
            self.file.write_group(key, host, {"time": np.float64(t),
                                              "step": np.int64(int(step))})
            self.file.flush()

        self._pending.append(self._worker.submit(job))

    def wait(self):
        """Block until every save_async has been written."""
        while self._pending:
            self._pending.pop(0).result()

    def list_snapshots(self):
        self.wait()
        return self.file.keys()

    def _prim(self, U, aux):
        if self.sim is None:
            raise ValueError("load needs the Simulator (H5Saver(..., simulator=sim))")
        import torch
        sim = self.sim
        Ut = torch.as_tensor(U, dtype=sim.dtype, device=sim.device)
        at = torch.as_tensor(aux, dtype=sim.dtype, device=sim.device)
        at = sim.update_func(Ut, at)
        rho = Ut[0]
        u = Ut[1] / rho
        v = Ut[2] / rho
        Y = Ut[4:] / rho
        inv_M = torch.as_tensor(1.0 / sim.mech.species_M, dtype=sim.dtype, device=sim.device)
        Y9 = torch.cat([Y, (1.0 - Y.sum(0))[None]])
        R = (inv_M[:, None, None] * Y9).sum(0)
        p = rho * R * at[1]
        return {"density": rho, "x-velocity": u, "y-velocity": v, "pressure": p,
                "tempreature": at[1], "species": Y}

    def load(self, key):
        """set_simulator.py:247-264: (time, primitives) of one snapshot, the
        temperature solved from the initial guess gamma = 1.4, T = 500 as
        the reference does (keys as the reference's, 'tempreature' included)."""
        self.wait()
        data, meta = self.file.read_group(key)
        U = data["u"]
        aux = np.empty((2,) + U.shape[1:], dtype=U.dtype)
        aux[0] = 1.4
        aux[1] = 500.0
        return float(meta["time"]), self._prim(U, aux)

    def load_all(self):
        """set_simulator.py:266-313: times and primitives of every snapshot,
        stacked on a leading axis in save_step order."""
        self.wait()

        def num(s):
            return int("".join(ch for ch in s if ch.isdigit()))

        import torch
        times, prims = [], []
        for key in sorted(self.file.keys(), key=num):
            t, q = self.load(key)
            times.append(t)
            prims.append(q)
        out = {k: torch.stack([q[k] for q in prims]) for k in prims[0]}
        return np.asarray(times), out

    def load_state(self, key):
        """(time, step, U, aux) of one snapshot for a restart: aux from the
        snapshot's 'aux' dataset when saved, else solved from the reference's
        initial guess."""
        self.wait()
        data, meta = self.file.read_group(key)
        U = data["u"]
        if "aux" in data:
            aux = data["aux"]
        else:
            if self.sim is None:
                raise ValueError("snapshot has no 'aux'; give the Simulator")
            import torch
            a0 = np.empty((2,) + U.shape[1:], dtype=U.dtype)
            a0[0] = 1.4
            a0[1] = 500.0
            Ut = torch.as_tensor(U, dtype=self.sim.dtype, device=self.sim.device)
            at = torch.as_tensor(a0, dtype=self.sim.dtype, device=self.sim.device)
            for _ in range(8):       # 5 Halley iterations per call from 500 K
                at = self.sim.update_func(Ut, at)
            aux = at
        return float(meta["time"]), int(meta["step"]), U, aux

    def close(self):
        self.wait()
        self.file.close()
        self._worker.shutdown()


def _to_host(v):
    if hasattr(v, "detach"):
        return v.detach().cpu().numpy()
    return np.asarray(v)
