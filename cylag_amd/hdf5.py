"""HDF5 files of named arrays, without h5py (triangular_mesh.pyx:204-286,
eulerian_field_set.pyx:486-604 read and write them with h5py).

CyLag's mesh and field-set files are flat: one root group holding named
datasets (numeric arrays, and byte strings for the optional field names).
This module reads and writes that subset of the HDF5 file format (version 0
superblock, version 1 object headers, groups indexed by a version 1 B-tree
and a local heap):

* H5Reader reads what h5py writes with its default settings: fixed-point,
  floating-point, fixed-length and variable-length string datasets with
  contiguous or compact layout, in groups stored as symbol tables or as
  compact link messages.  Chunked or filtered (compressed) datasets and
  dense link storage raise NotImplementedError.
* write_h5 writes the same structures: every dataset contiguous, strings
  fixed-length (h5py and the reference's readers take them as bytes).

Byte layouts follow the HDF5 File Format Specification, version 2 document
(superblock, symbol table entry and node, v1 B-tree, local and global heaps,
object header v1, dataspace, datatype, data layout, link and symbol table
messages).
"""
import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF

# object header message types
MSG_NIL, MSG_DATASPACE, MSG_LINK_INFO, MSG_DATATYPE = 0x0, 0x1, 0x2, 0x3
MSG_LINK, MSG_LAYOUT, MSG_FILTERS, MSG_CONT, MSG_SYMTAB = 0x6, 0x8, 0xB, 0x10, 0x11


class H5Reader:
    """Read-only view of an HDF5 file's root group: `keys()`, `in`, and
    `reader[name]` returning a numpy array (h5py's `f[name][:]`)."""

    def __init__(self, file_name):
        with open(file_name, "rb") as f:
            self.buf = f.read()
        b = self.buf
        if b[:8] != SIGNATURE:
            raise ValueError("%s: not an HDF5 file" % file_name)
        ver = b[8]
        if ver not in (0, 1):
            raise NotImplementedError("HDF5 superblock version %d (only 0 and 1 are read)" % ver)
        self.O, self.L = b[13], b[14]
        p = 24 if ver == 0 else 28
        self.base = self._addr(p)
        root_entry = p + 4 * self.O
        root_header = self._addr(root_entry + self.O)
        self.members = self._group(root_header)

    # -- primitive reads ------------------------------------------------------------
    def _uint(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _addr(self, p):
        return self._uint(p, self.O)

    def _at(self, a):
        return self.base + a

    def _cstr(self, p):
        e = self.buf.index(b"\0", p)
        return self.buf[p:e].decode("utf-8")

    # -- object headers ---------------------------------------------------------------
    def _messages(self, a):
        """(type, data offset, size) of every message of the v1 object header at a."""
        b, p = self.buf, self._at(a)
        if b[p] != 1:
            raise NotImplementedError("object header version %d (only version 1 is read)" % b[p])
        blocks = [(p + 16, self._uint(p + 8, 4))]
        out = []
        while blocks:
            q, n = blocks.pop(0)
            end = q + n
            while q + 8 <= end:
                t, size = self._uint(q, 2), self._uint(q + 2, 2)
                d = q + 8
                if t == MSG_CONT:
                    blocks.append((self._at(self._addr(d)), self._uint(d + self.O, self.L)))
                elif t != MSG_NIL:
                    out.append((t, d, size))
                q = d + size
        return out

    def _group(self, a):
        """{name: object header address} of the group whose header is at a."""
        members = {}
        for t, d, _ in self._messages(a):
            if t == MSG_SYMTAB:
                btree, heap = self._addr(d), self._addr(d + self.O)
                heap_data = self._heap_data(heap)
                self._btree_group(btree, heap_data, members)
            elif t == MSG_LINK:
                name, target = self._link(d)
                if target is not None:
                    members[name] = target
            elif t == MSG_LINK_INFO:
                fheap = self._addr(d + 2 + (8 if self.buf[d + 1] & 1 else 0))
                if fheap != UNDEF:
                    raise NotImplementedError("dense link storage (fractal heap) is not read")
        return members

    def _heap_data(self, a):
        p = self._at(a)
        if self.buf[p:p + 4] != b"HEAP":
            raise ValueError("bad local heap signature at %d" % p)
        return self._at(self._addr(p + 8 + 2 * self.L))

    def _btree_group(self, a, heap_data, members):
        p = self._at(a)
        b = self.buf
        if b[p:p + 4] != b"TREE" or b[p + 4] != 0:
            raise ValueError("bad group B-tree node at %d" % p)
        level, used = b[p + 5], self._uint(p + 6, 2)
        q = p + 8 + 2 * self.O + self.L  # first child (after key 0)
        for _ in range(used):
            child = self._addr(q)
            if level > 0:
                self._btree_group(child, heap_data, members)
            else:
                self._snod(child, heap_data, members)
            q += self.O + self.L

    def _snod(self, a, heap_data, members):
        p = self._at(a)
        if self.buf[p:p + 4] != b"SNOD":
            raise ValueError("bad symbol table node at %d" % p)
        n = self._uint(p + 6, 2)
        q = p + 8
        for _ in range(n):
            name = self._cstr(heap_data + self._uint(q, self.O))
            members[name] = self._addr(q + self.O)
            q += 2 * self.O + 24

    def _link(self, d):
        b = self.buf
        flags = b[d + 1]
        q = d + 2
        ltype = 0
        if flags & 0x8:
            ltype = b[q]
            q += 1
        if flags & 0x4:
            q += 8
        if flags & 0x10:
            q += 1
        nl = 1 << (flags & 3)
        n = self._uint(q, nl)
        q += nl
        name = b[q:q + n].decode("utf-8")
        q += n
        return name, (self._addr(q) if ltype == 0 else None)

    # -- datasets ---------------------------------------------------------------------------
    def keys(self):
        return list(self.members)

    def __contains__(self, name):
        return name in self.members

    def __getitem__(self, name):
        return self.read(name)

    def read(self, name):
        if name not in self.members:
            raise KeyError(name)
        shape = dtype = layout = vlen = None
        for t, d, size in self._messages(self.members[name]):
            if t == MSG_DATASPACE:
                shape = self._dataspace(d)
            elif t == MSG_DATATYPE:
                dtype, vlen = self._datatype(d)
            elif t == MSG_LAYOUT:
                layout = self._layout(d)
            elif t == MSG_FILTERS:
                raise NotImplementedError("%s: filtered (compressed) datasets are not read" % name)
        if shape is None or dtype is None or layout is None:
            raise ValueError("%s is not a dataset" % name)
        count = int(np.prod(shape, dtype=np.int64))
        if vlen:
            raw = self._raw(layout, count * (4 + self.O + 4))
            out = np.empty(count, dtype=object)
            for i in range(count):
                r = i * (8 + self.O)
                ln = int.from_bytes(raw[r:r + 4], "little")
                coll = int.from_bytes(raw[r + 4:r + 4 + self.O], "little")
                idx = int.from_bytes(raw[r + 4 + self.O:r + 8 + self.O], "little")
                out[i] = self._global_object(coll, idx)[:ln]
            return out.reshape(shape)
        raw = self._raw(layout, count * dtype.itemsize)
        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).copy()

    def _dataspace(self, d):
        b = self.buf
        ver, nd = b[d], b[d + 1]
        if ver == 1:
            q = d + 8
        elif ver == 2:
            if b[d + 3] == 2:
                return (0,)
            q = d + 4
        else:
            raise NotImplementedError("dataspace message version %d" % ver)
        return tuple(self._uint(q + i * self.L, self.L) for i in range(nd))

    def _datatype(self, d):
        b = self.buf
        cls = b[d] & 0x0F
        bits0 = b[d + 1]
        size = self._uint(d + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 0:
            return np.dtype("%s%s%d" % (order, "i" if bits0 & 0x8 else "u", size)), False
        if cls == 1:
            if size not in (2, 4, 8):
                raise NotImplementedError("floating-point size %d" % size)
            return np.dtype("%sf%d" % (order, size)), False
        if cls == 3:
            return np.dtype("S%d" % size), False
        if cls == 9 and (bits0 & 0x0F) == 1:
            return np.dtype(object), True
        raise NotImplementedError("HDF5 datatype class %d" % cls)

    def _layout(self, d):
        b = self.buf
        ver = b[d]
        if ver == 3:
            cls = b[d + 1]
            if cls == 0:
                n = self._uint(d + 2, 2)
                return ("compact", d + 4, n)
            if cls == 1:
                return ("contiguous", self._addr(d + 2), self._uint(d + 2 + self.O, self.L))
            raise NotImplementedError("chunked datasets are not read")
        if ver in (1, 2):
            nd, cls = b[d + 1], b[d + 2]
            q = d + 8
            if cls == 0:
                q += 4 * nd
                return ("compact", q + 4, self._uint(q, 4))
            if cls == 1:
                return ("contiguous", self._addr(q), None)
            raise NotImplementedError("chunked datasets are not read")
        raise NotImplementedError("data layout message version %d" % ver)

    def _raw(self, layout, nbytes):
        kind, a, n = layout
        if kind == "compact":
            return self.buf[a:a + nbytes]
        if a == UNDEF or nbytes == 0:
            return bytes(nbytes)  # storage never written: zero fill
        p = self._at(a)
        return self.buf[p:p + nbytes]

    def _global_object(self, coll, idx):
        p = self._at(coll)
        if self.buf[p:p + 4] != b"GCOL":
            raise ValueError("bad global heap collection at %d" % p)
        end = p + self._uint(p + 8, self.L)
        q = p + 8 + self.L
        while q + 8 + self.L <= end:
            i = self._uint(q, 2)
            n = self._uint(q + 8, self.L)
            if i == 0:
                break
            if i == idx:
                return self.buf[q + 8 + self.L:q + 8 + self.L + n]
            q += 8 + self.L + ((n + 7) & ~7)
        raise ValueError("global heap object %d not found" % idx)


# -- writer ------------------------------------------------------------------------------------
LEAF_K = 4        # symbols per symbol table node: 2 * LEAF_K
NODE_K = 16       # children per group B-tree node: 2 * NODE_K


def _pad8(b):
    return b + bytes(-len(b) % 8)


def _message(t, data):
    data = _pad8(data)
    return struct.pack("<HHB3x", t, len(data), 0) + data


def _object_header(messages):
    body = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


def _datatype_message(dt):
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0
        return struct.pack("<B3BIHH", 0x10, bits, 0, 0, dt.itemsize, 0, 8 * dt.itemsize)
    if dt.kind == "f":
        if dt.itemsize == 8:
            sign, exp_loc, exp_size, mant, bias = 63, 52, 11, 52, 1023
        elif dt.itemsize == 4:
            sign, exp_loc, exp_size, mant, bias = 31, 23, 8, 23, 127
        else:
            raise NotImplementedError("float%d" % (8 * dt.itemsize))
        return struct.pack("<B3BIHHBBBBI", 0x11, 0x20, sign, 0, dt.itemsize, 0, 8 * dt.itemsize,
                           exp_loc, exp_size, 0, mant, bias)
    if dt.kind == "S":
        return struct.pack("<B3BI", 0x13, 1, 0, 0, max(dt.itemsize, 1))  # null-padded ASCII
    raise NotImplementedError("dtype %s" % dt)


def _as_dataset(value):
    if isinstance(value, (list, tuple)) and value and isinstance(value[0], (str, bytes)):
        value = [v.encode("utf-8") if isinstance(v, str) else v for v in value]
        return np.array(value, dtype="S%d" % max(1, max(len(v) for v in value)))
    a = np.asarray(value)
    if a.dtype.kind == "U":
        a = np.char.encode(a, "utf-8")
    if a.dtype.kind == "b":
        a = a.astype(np.uint8)
    if a.dtype.kind not in "iufS":
        raise NotImplementedError("dtype %s" % a.dtype)
    if a.dtype.byteorder == ">":
        a = a.astype(a.dtype.newbyteorder("<"))
    return np.ascontiguousarray(a)


def write_h5(file_name, datasets):
    """Write {name: array} as datasets of the root group of a new HDF5 file.
    Arrays are stored contiguous in C order, little-endian; lists of str or
    bytes become fixed-length byte strings."""
    names = sorted(datasets, key=lambda n: n.encode("utf-8"))
    arrays = [_as_dataset(datasets[n]) for n in names]
    O = 8
    # local heap: offset 0 holds the empty name, then each member name
    heap = bytearray(8)
    name_off = []
    for n in names:
        name_off.append(len(heap))
        heap += _pad8(n.encode("utf-8") + b"\0")
    # one free block closes the data segment, as the HDF5 library writes it:
    # (offset of the next free block, 1 = none; size of this block)
    free_off = len(heap)
    heap += struct.pack("<QQ", 1, 16)
    snods = [names[i:i + 2 * LEAF_K] for i in range(0, len(names), 2 * LEAF_K)] or [[]]
    if len(snods) > 2 * NODE_K:
        raise NotImplementedError("more than %d datasets in one file" % (2 * NODE_K * 2 * LEAF_K))
    snod_size = 8 + 2 * LEAF_K * 40
    btree_size = 24 + (2 * NODE_K + 1) * 8 + 2 * NODE_K * O

    out = bytearray()
    sb_size = 24 + 4 * O + 40

    def alloc(n):
        a = sb_size + len(out)
        out.extend(bytes(n))
        return a

    def put(a, data):
        out[a - sb_size:a - sb_size + len(data)] = data

    root_hdr = alloc(len(_object_header([_message(MSG_SYMTAB, bytes(2 * O))])))
    heap_hdr = alloc(32)
    heap_data = alloc(len(heap))
    btree = alloc(btree_size)
    snod_addr = [alloc(snod_size) for _ in snods]
    obj_addr, data_addr = [], []
    for a in arrays:
        space = struct.pack("<BBB5x", 1, a.ndim, 0) + b"".join(struct.pack("<Q", d) for d in a.shape)
        msgs = [_message(MSG_DATASPACE, space), _message(MSG_DATATYPE, _datatype_message(a.dtype)),
                _message(MSG_LAYOUT, struct.pack("<BBQQ", 3, 1, 0, 0))]
        obj_addr.append(alloc(len(_object_header(msgs))))
    for a in arrays:
        data_addr.append(alloc(len(_pad8(a.tobytes()))) if a.nbytes else UNDEF)

    put(root_hdr, _object_header([_message(MSG_SYMTAB, struct.pack("<QQ", btree, heap_hdr))]))
    put(heap_hdr, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), free_off, heap_data))
    put(heap_data, bytes(heap))
    node = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), UNDEF, UNDEF))
    node += struct.pack("<Q", 0)
    k = 0
    for i, group in enumerate(snods):
        k += len(group)
        node += struct.pack("<QQ", snod_addr[i], name_off[k - 1] if group else 0)
    put(btree, bytes(node))
    k = 0
    for i, group in enumerate(snods):
        rec = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(group)))
        for _ in group:
            rec += struct.pack("<QQII16x", name_off[k], obj_addr[k], 0, 0)
            k += 1
        put(snod_addr[i], bytes(rec))
    for a, oa, da in zip(arrays, obj_addr, data_addr):
        space = struct.pack("<BBB5x", 1, a.ndim, 0) + b"".join(struct.pack("<Q", d) for d in a.shape)
        msgs = [_message(MSG_DATASPACE, space), _message(MSG_DATATYPE, _datatype_message(a.dtype)),
                _message(MSG_LAYOUT, struct.pack("<BBQQ", 3, 1, da, a.nbytes))]
        put(oa, _object_header(msgs))
        if a.nbytes:
            put(da, a.tobytes())
    eof = sb_size + len(out)
    sb = SIGNATURE + struct.pack("<8B", 0, 0, 0, 0, 0, O, 8, 0) + struct.pack("<HHI", LEAF_K, NODE_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", btree, heap_hdr)
    assert len(sb) == sb_size
    with open(file_name, "wb") as f:
        f.write(sb)
        f.write(out)
