"""Minimal read-only HDF5 reader, used where h5py is not available.

Covers what the reference's data files use: superblock v0/v1, version-1 object
headers (with continuation blocks), old-style groups (symbol-table B-tree v1 +
local heap), and contiguous or compact datasets of little/big-endian integers
and IEEE floats with simple dataspaces.  Anything else raises
NotImplementedError.  Reading follows the published HDF5 File Format
Specification (version 2.0), sections II (superblock), III.A-C (B-trees,
symbol table nodes, local heaps) and IV.A (object headers and messages).
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
        return arr.astype(dtype.newbyteorder("="))

    def _dataspace(self, d):
        ver = self.buf[d]
        rank = self.buf[d + 1]
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
