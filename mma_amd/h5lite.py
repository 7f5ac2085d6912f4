"""Minimal HDF5 reader and writer, used where h5py is not available.

Covers what the reference's data files use: superblock v0/v1, version-1 object
headers (with continuation blocks), old-style groups (symbol-table B-tree v1 +
local heap), contiguous or compact datasets of little/big-endian integers, IEEE
floats and fixed-length strings with scalar or simple dataspaces, and
attributes of those types.  Anything else raises NotImplementedError.  Both
directions follow the published HDF5 File Format Specification (version 2.0),
sections II (superblock), III.A-C (B-trees, symbol table nodes, local heaps) and
IV.A (object headers and messages); `write` produces files of exactly that subset
(superblock v0, 8-byte offsets and lengths, contiguous datasets), which the HDF5
library reads as it reads files written by h5py with default settings.
"""
import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class H5File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        if self.buf[:8] != SIGNATURE:
            raise NotImplementedError("superblock not at offset 0")
        ver = self.buf[8]
        if ver not in (0, 1):
            raise NotImplementedError("superblock version %d" % ver)
        self.so = self.buf[13]  # size of offsets
        self.sl = self.buf[14]  # size of lengths
        p = 24 if ver == 0 else 28
        self.base = self._off(p)
        p += 4 * self.so
        # root group symbol table entry
        self.root = self._entry(p)

    # -- primitive readers -------------------------------------------------
    def _uint(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _off(self, p):
        return self._uint(p, self.so)

    def _len(self, p):
        return self._uint(p, self.sl)

    def _entry(self, p):
        name_off = self._off(p)
        hdr = self._off(p + self.so)
        cache = self._uint(p + 2 * self.so, 4)
        scratch = p + 2 * self.so + 8
        btree = heap = None
        if cache == 1:
            btree = self._off(scratch)
            heap = self._off(scratch + self.so)
        return {"name_off": name_off, "header": hdr, "btree": btree, "heap": heap}

    # -- object headers ----------------------------------------------------
    def _messages(self, addr):
        a = self.base + addr
        if self.buf[a] != 1:
            raise NotImplementedError("object header version %d" % self.buf[a])
        nmsg = self._uint(a + 2, 2)
        size = self._uint(a + 8, 4)
        blocks = [(a + 16, size)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(msgs) < nmsg:
                mtype = self._uint(p, 2)
                msize = self._uint(p + 2, 2)
                data = p + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self.base + self._off(data), self._len(data + self.so)))
                msgs.append((mtype, data, msize))
                p = data + msize
        return msgs

    def _heap_name(self, heap_addr, off):
        h = self.base + heap_addr
        if self.buf[h:h + 4] != b"HEAP":
            raise NotImplementedError("bad local heap")
        data = self.base + self._off(h + 8 + 2 * self.sl)
        end = self.buf.index(b"\x00", data + off)
        return self.buf[data + off:end].decode()

    def _group_children(self, btree, heap):
        out = {}
        stack = [btree]
        while stack:
            n = self.base + stack.pop()
            if self.buf[n:n + 4] != b"TREE":
                raise NotImplementedError("bad B-tree node")
            level = self.buf[n + 5]
            used = self._uint(n + 6, 2)
            p = n + 8 + 2 * self.so
            for _ in range(used):
                p += self.sl  # key
                child = self._off(p)
                p += self.so
                if level > 0:
                    stack.append(child)
                else:
                    s = self.base + child
                    if self.buf[s:s + 4] != b"SNOD":
                        raise NotImplementedError("bad symbol node")
                    nsym = self._uint(s + 6, 2)
                    q = s + 8
                    for _ in range(nsym):
                        e = self._entry(q)
                        out[self._heap_name(heap, e["name_off"])] = e
                        q += 2 * self.so + 24
        return out

    def _group_of(self, entry):
        if entry["btree"] is not None:
            return entry["btree"], entry["heap"]
        for mtype, data, _ in self._messages(entry["header"]):
            if mtype == 0x11:
                return self._off(data), self._off(data + self.so)
        return None

    # -- public API --------------------------------------------------------
    def keys(self, path="/"):
        entry = self._lookup(path)
        g = self._group_of(entry)
        if g is None:
            raise KeyError(path + " is not a group")
        return sorted(self._group_children(*g).keys())

    def _lookup(self, path):
        entry = self.root
        for part in [s for s in path.split("/") if s]:
            g = self._group_of(entry)
            if g is None:
                raise KeyError(path)
            children = self._group_children(*g)
            if part not in children:
                raise KeyError(path)
            entry = children[part]
        return entry

    def __contains__(self, path):
        try:
            self._lookup(path)
        except KeyError:
            return False
        return True

    def is_group(self, path):
        return self._group_of(self._lookup(path)) is not None

    def __getitem__(self, path):
        entry = self._lookup(path)
        dims = dtype = None
        layout = None
        for mtype, d, msize in self._messages(entry["header"]):
            if mtype == 0x01:
                dims = self._dataspace(d)
            elif mtype == 0x03:
                dtype = self._datatype(d)
            elif mtype == 0x08:
                layout = (d, msize)
        if dims is None or dtype is None or layout is None:
            raise KeyError(path + " is not a dataset")
        raw = self._layout(layout[0], int(np.prod(dims, dtype=np.int64)) * dtype.itemsize)
        arr = np.frombuffer(raw, dtype=dtype).reshape(dims)
        return arr.astype(dtype.newbyteorder("=")) if dtype.kind != "S" else arr.copy()

    def attrs(self, path):
        """The attributes of an object: {name: value}, strings as bytes, numbers as arrays
        (0-d for scalars)."""
        out = {}
        for mtype, d, msize in self._messages(self._lookup(path)["header"]):
            if mtype != 0x0C:
                continue
            ver = self.buf[d]
            name_size, dt_size, ds_size = self._uint(d + 2, 2), self._uint(d + 4, 2), self._uint(d + 6, 2)
            if ver == 1:
                pad = lambda n: (n + 7) // 8 * 8  # noqa: E731
                p = d + 8
                name = self.buf[p:p + name_size].split(b"\x00")[0].decode()
                p += pad(name_size)
                dtype = self._datatype(p)
                p += pad(dt_size)
                dims = self._dataspace(p)
                p += pad(ds_size)
            elif ver in (2, 3):
                p = d + 8 + (1 if ver == 3 else 0)
                name = self.buf[p:p + name_size].split(b"\x00")[0].decode()
                p += name_size
                dtype = self._datatype(p)
                p += dt_size
                dims = self._dataspace(p)
                p += ds_size
            else:
                raise NotImplementedError("attribute message version %d" % ver)
            n = int(np.prod(dims, dtype=np.int64)) * dtype.itemsize
            arr = np.frombuffer(self.buf[p:p + n], dtype=dtype).reshape(dims)
            out[name] = arr.tobytes() if dtype.kind == "S" and arr.ndim == 0 else arr.copy()
        return out

    def _dataspace(self, d):
        ver = self.buf[d]
        rank = self.buf[d + 1]
        if ver == 2 and self.buf[d + 3] == 0:
            return ()  # scalar
        p = d + 8 if ver == 1 else d + 4
        return tuple(self._len(p + i * self.sl) for i in range(rank))

    def _datatype(self, d):
        cls = self.buf[d] & 0x0F
        bits = self.buf[d + 1]
        size = self._uint(d + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            signed = bool(bits & 0x08)
            return np.dtype(order + ("i" if signed else "u") + str(size))
        if cls == 1:
            return np.dtype(order + "f" + str(size))
        if cls == 3:
            return np.dtype("S%d" % size)
        raise NotImplementedError("datatype class %d" % cls)

    def _layout(self, d, nbytes):
        ver = self.buf[d]
        if ver == 3:
            cls = self.buf[d + 1]
            if cls == 0:  # compact
                size = self._uint(d + 2, 2)
                return self.buf[d + 4:d + 4 + size][:nbytes]
            if cls == 1:
                addr = self._off(d + 2)
                if addr == (1 << (8 * self.so)) - 1:
                    return bytes(nbytes)  # never written: fill value 0
                return self.buf[self.base + addr:self.base + addr + nbytes]
            raise NotImplementedError("chunked layout")
        if ver in (1, 2):
            cls = self.buf[d + 2]
            if cls == 1:
                addr = self._off(d + 8)
                return self.buf[self.base + addr:self.base + addr + nbytes]
            raise NotImplementedError("layout class %d" % cls)
        raise NotImplementedError("layout version %d" % ver)


# ---------------------------------------------------------------------------------------------
# Writer

UNDEF = (1 << 64) - 1   # undefined address (all bits set)
HEAP_FREE_NULL = 1      # local heap: no free block (the library's end-of-free-list value)
LEAF_K, INTERNAL_K = 4, 16  # symbol-table node and group B-tree node widths (the defaults)


def _pad8(b):
    return b + bytes((-len(b)) % 8)


def _dtype_message(dt):
    """Datatype message body (IV.A.2.d), version 1."""
    dt = np.dtype(dt)
    order = 1 if dt.byteorder == ">" else 0
    if dt.kind == "f":
        if dt.itemsize == 8:
            sign, eloc, esz, msz, bias = 63, 52, 11, 52, 1023
        elif dt.itemsize == 4:
            sign, eloc, esz, msz, bias = 31, 23, 8, 23, 127
        else:
            raise NotImplementedError("float of %d bytes" % dt.itemsize)
        # mantissa normalisation 2 (most significant bit implied), bits 4-5 of the first byte
        return (struct.pack("<BBBBI", 0x10 | 1, order | 0x20, sign, 0, dt.itemsize)
                + struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esz, 0, msz, bias))
    if dt.kind in "iu":
        signed = 0x08 if dt.kind == "i" else 0
        return struct.pack("<BBBBI", 0x10 | 0, order | signed, 0, 0, dt.itemsize) + struct.pack("<HH", 0,
                                                                                              8 * dt.itemsize)
    if dt.kind == "S":
        return struct.pack("<BBBBI", 0x10 | 3, 0x01, 0, 0, dt.itemsize)  # null-padded ASCII
    raise NotImplementedError("datatype %s" % dt)


def _dataspace_message(shape):
    """Dataspace message body (IV.A.2.b), version 1: rank 0 is a scalar."""
    return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", int(n)) for n in shape)


def _message(mtype, body, flags=0):
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _object_header(messages):
    """Version-1 object header (IV.A.1.a) holding the given messages."""
    data = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(data)) + data


def _attribute_message(name, value):
    """Attribute message body (IV.A.2.m), version 1; value: bytes/str (fixed-length string) or a
    number / array."""
    if isinstance(value, str):
        value = value.encode()
    if isinstance(value, bytes):
        arr = np.array(value, dtype="S%d" % max(1, len(value)))
    else:
        arr = np.asarray(value, order="C")
        if arr.dtype.kind not in "fiuS":
            raise NotImplementedError("attribute of type %s" % arr.dtype)
        arr = arr.astype(arr.dtype.newbyteorder("<")) if arr.dtype.kind != "S" else arr
    nm = name.encode() + b"\x00"
    dt = _dtype_message(arr.dtype)
    ds = _dataspace_message(arr.shape)
    return (struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds)) + _pad8(nm) + _pad8(dt) + _pad8(ds)
            + arr.tobytes())


def _as_array(value):
    if isinstance(value, str):
        value = value.encode()
    if isinstance(value, bytes):
        return np.array(value, dtype="S%d" % max(1, len(value)))
    a = np.asarray(value)
    if a.dtype.kind == "c":
        raise NotImplementedError("complex datasets: store (..., 2) re/im arrays")
    if a.dtype.kind not in "fiuS":
        raise NotImplementedError("dataset of type %s" % a.dtype)
    if a.dtype.kind != "S" and a.dtype.byteorder == ">":
        a = a.astype(a.dtype.newbyteorder("<"))
    return np.asarray(a, order="C")  # (ascontiguousarray would turn a scalar into shape (1,))


def write(path, datasets, attrs=None):
    """Writes an HDF5 file holding `datasets` {"/a/b/name": array-like} (intermediate groups
    created; numbers as given, bytes/str as fixed-length strings, 0-d arrays and Python scalars as
    scalar datasets) and `attrs` {"/a/b/name": {attribute: value}} on datasets or groups.  The
    array data is streamed to the file first; the metadata (object headers, group B-trees, symbol
    table nodes, local heaps) follows it, and the superblock at offset 0 points to the root group."""
    attrs = attrs or {}
    tree = {}  # nested dict of groups; leaves are ("ds", key)
    items = {}
    for key, value in datasets.items():
        parts = [p for p in key.split("/") if p]
        if not parts:
            raise ValueError("empty dataset path")
        node = tree
        for p in parts[:-1]:
            node = node.setdefault(p, {})
            if not isinstance(node, dict):
                raise ValueError("%s: %s is a dataset" % (key, p))
        if parts[-1] in node:
            raise ValueError("%s given twice" % key)
        node[parts[-1]] = ("ds", "/" + "/".join(parts))
        items["/" + "/".join(parts)] = _as_array(value)
    norm_attrs = {"/" + "/".join(p for p in k.split("/") if p): v for k, v in attrs.items()}

    with open(path, "wb") as fh:
        fh.write(bytes(96))  # superblock, written last
        data_addr = {}
        for key, arr in items.items():
            pos = fh.tell()
            if pos % 8:
                fh.write(bytes(8 - pos % 8))
                pos = fh.tell()
            data_addr[key] = pos
            if arr.nbytes:
                fh.write(memoryview(arr.reshape(-1)).cast("B"))
        pos = fh.tell()
        if pos % 8:
            fh.write(bytes(8 - pos % 8))
        meta = bytearray()
        base = fh.tell()

        def put(blob):
            addr = base + len(meta)
            meta.extend(blob)
            return addr

        def attr_messages(key):
            return [_message(0x0C, _attribute_message(n, v)) for n, v in norm_attrs.get(key, {}).items()]

        def dataset_header(key):
            arr = items[key]
            msgs = [_message(0x01, _dataspace_message(arr.shape)),
                    _message(0x03, _dtype_message(arr.dtype), flags=1),
                    # fill value (IV.A.2.f) version 2: late allocation, written if set, none defined
                    _message(0x05, struct.pack("<BBBB", 2, 2, 2, 0)),
                    # layout (IV.A.2.i) version 3, contiguous
                    _message(0x08, struct.pack("<BBQQ", 3, 1, data_addr[key] if arr.nbytes else UNDEF, arr.nbytes))]
            return put(_object_header(msgs + attr_messages(key)))

        def group(node, key):
            """Writes a group's children, heap, symbol nodes and B-tree; returns (header, btree, heap)."""
            names = sorted(node, key=lambda n: n.encode())
            entries = []
            for n in names:
                child = node[n]
                ck = (key.rstrip("/") + "/" + n)
                if isinstance(child, dict):
                    h, bt, hp = group(child, ck)
                    entries.append((n, h, (bt, hp)))
                else:
                    entries.append((n, dataset_header(ck), None))
            # local heap: the empty string at offset 0, then the names, each 8-byte aligned
            heap_data = bytearray(8)
            name_off = {}
            for n in names:
                name_off[n] = len(heap_data)
                heap_data.extend(_pad8(n.encode() + b"\x00"))
            heap_data_addr = put(bytes(heap_data))
            heap = put(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), HEAP_FREE_NULL, heap_data_addr))
            # symbol table nodes of up to 2K entries (III.B), full-size
            snods = []
            for i in range(0, max(len(entries), 1), 2 * LEAF_K):
                chunk = entries[i:i + 2 * LEAF_K]
                body = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(chunk)))
                for n, h, grp in chunk:
                    if grp is None:
                        body.extend(struct.pack("<QQI4x16x", name_off[n], h, 0))
                    else:
                        body.extend(struct.pack("<QQI4xQQ", name_off[n], h, 1, grp[0], grp[1]))
                body.extend(bytes((2 * LEAF_K - len(chunk)) * 40))
                snods.append((put(bytes(body)), chunk[-1][0] if chunk else None))
            if len(snods) > 2 * INTERNAL_K:
                raise NotImplementedError("more than %d children in one group" % (2 * INTERNAL_K * 2 * LEAF_K))
            # one leaf B-tree node (III.A.1) over the symbol nodes: key 0 is the empty string, key
            # i+1 the last name of symbol node i
            bt = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods) if entries else 0, UNDEF, UNDEF))
            bt.extend(struct.pack("<Q", 0))
            for addr, last in snods if entries else []:
                bt.extend(struct.pack("<QQ", addr, name_off[last]))
            bt.extend(bytes(24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8 - len(bt)))
            btree = put(bytes(bt))
            header = put(_object_header([_message(0x11, struct.pack("<QQ", btree, heap))] + attr_messages(key)))
            return header, btree, heap

        root_header, root_btree, root_heap = group(tree, "/")
        fh.write(bytes(meta))
        eof = fh.tell()
        sb = (SIGNATURE + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, LEAF_K, INTERNAL_K, 0)
              + struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
              + struct.pack("<QQI4xQQ", 0, root_header, 1, root_btree, root_heap))
        assert len(sb) == 96
        fh.seek(0)
        fh.write(sb)


def read_all(path):
    """Every dataset of a file as {"/a/b/name": array} and the attributes as {path: {name: value}}
    (groups included where they carry attributes)."""
    f = H5File(path)
    data, attrs = {}, {}

    def walk(p):
        a = f.attrs(p)
        if a:
            attrs[p] = a
        if f.is_group(p):
            for k in f.keys(p):
                walk(p.rstrip("/") + "/" + k)
        else:
            data[p] = f[p]

    walk("/")
    return data, attrs
