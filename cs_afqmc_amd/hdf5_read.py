"""Minimal HDF5 reader for FCIDUMP_chol files (no h5py needed).

pyscf_interface.py:578-598 of the reference writes FCIDUMP_chol with h5py: a handful of
root-group datasets (header, energy_core, hcore, hcore_mod, chol, mo_coeffs_up/dn) of
little-endian int64 / float64 in contiguous (or compact) storage, in a file with a
version-0 superblock, a symbol-table root group (version-1 B-tree of symbol nodes, local
heap of names) and version-1 object headers.  This module reads exactly that subset of
the HDF5 file format specification; anything else (chunked or filtered storage, other
superblock or object-header versions, nested groups, non-numeric types) raises
ValueError.
"""

import struct

import numpy as np


class _File:
    def __init__(self, data):
        self.d = data

    def u(self, off, size):
        if off < 0 or off + size > len(self.d):
            raise ValueError("HDF5: read beyond the end of the file")
        return int.from_bytes(self.d[off:off + size], "little")

    def sig(self, off, s):
        if self.d[off:off + len(s)] != s:
            raise ValueError(f"HDF5: expected {s!r} at {off}")


def read_datasets(path):
    """{name: numpy array} for every dataset of the root group of the HDF5 file."""
    with open(path, "rb") as fh:
        f = _File(fh.read())
    f.sig(0, b"\x89HDF\r\n\x1a\n")
    if f.u(8, 1) != 0:
        raise ValueError(f"HDF5: superblock version {f.u(8, 1)} is not supported")
    so, sl = f.u(13, 1), f.u(14, 1)
    if so != 8 or sl != 8:
        raise ValueError("HDF5: only 8-byte offsets and lengths are supported")
    base = f.u(24, 8)
    root = 56  # root group symbol table entry
    if f.u(root + 16, 4) != 1:
        raise ValueError("HDF5: the root group has no cached symbol table")
    btree = f.u(root + 24, 8) + base
    heap = f.u(root + 32, 8) + base
    f.sig(heap, b"HEAP")
    heap_data = f.u(heap + 24, 8) + base

    def name_at(off):
        start = heap_data + off
        end = f.d.index(b"\x00", start)
        return f.d[start:end].decode("utf-8")

    entries = []

    def walk(node):
        f.sig(node, b"TREE")
        if f.u(node + 4, 1) != 0:
            raise ValueError("HDF5: not a group B-tree")
        level, used = f.u(node + 5, 1), f.u(node + 6, 2)
        p = node + 8 + 16 + 8  # signature/type/level/used, siblings, key 0
        for _ in range(used):
            child = f.u(p, 8) + base
            p += 16  # child address, next key
            if level > 0:
                walk(child)
            else:
                f.sig(child, b"SNOD")
                for k in range(f.u(child + 6, 2)):
                    e = child + 8 + 40 * k
                    entries.append((name_at(f.u(e, 8)), f.u(e + 8, 8) + base))

    walk(btree)
    return {name: _dataset(f, addr, base) for name, addr in entries}


def _messages(f, header, base):
    if f.u(header, 1) != 1:
        raise ValueError("HDF5: only version-1 object headers are supported")
    nmsg = f.u(header + 2, 2)
    blocks = [(header + 16, f.u(header + 8, 4))]
    out = []
    while blocks and len(out) < nmsg:
        p, size = blocks.pop(0)
        end = p + size
        while p + 8 <= end and len(out) < nmsg:
            mtype, msize = f.u(p, 2), f.u(p + 2, 2)
            body = p + 8
            if mtype == 0x10:  # continuation
                blocks.append((f.u(body, 8) + base, f.u(body + 8, 8)))
            out.append((mtype, body, msize))
            p = body + msize
    return out


def _dataset(f, header, base):
    shape = dtype = None
    layout = None
    for mtype, p, size in _messages(f, header, base):
        if mtype == 0x0001:  # dataspace
            ver, rank, flags = f.u(p, 1), f.u(p + 1, 1), f.u(p + 2, 1)
            if ver == 1:
                dims_at = p + 8
            elif ver == 2:
                dims_at = p + 4
                if f.u(p + 3, 1) == 0:
                    rank = 0
            else:
                raise ValueError(f"HDF5: dataspace version {ver}")
            shape = tuple(f.u(dims_at + 8 * i, 8) for i in range(rank))
        elif mtype == 0x0003:  # datatype
            cls, bits, esize = f.u(p, 1) & 0x0F, f.u(p + 1, 1), f.u(p + 4, 4)
            if bits & 1:
                raise ValueError("HDF5: big-endian data is not supported")
            if cls == 0:
                dtype = np.dtype(f"<{'i' if bits & 0x08 else 'u'}{esize}")
            elif cls == 1 and esize in (4, 8):
                dtype = np.dtype(f"<f{esize}")
            else:
                raise ValueError(f"HDF5: datatype class {cls} is not supported")
        elif mtype == 0x0008:  # data layout
            ver = f.u(p, 1)
            if ver != 3:
                raise ValueError(f"HDF5: data layout version {ver} is not supported")
            cls = f.u(p + 1, 1)
            if cls == 1:  # contiguous
                layout = ("contiguous", f.u(p + 2, 8), f.u(p + 10, 8))
            elif cls == 0:  # compact
                n = f.u(p + 2, 2)
                layout = ("compact", p + 4, n)
            else:
                raise ValueError("HDF5: chunked storage is not supported")
        elif mtype == 0x000B:  # filter pipeline
            raise ValueError("HDF5: filtered datasets are not supported")
    if shape is None or dtype is None or layout is None:
        raise ValueError("HDF5: object is not a dataset")
    count = int(np.prod(shape)) if shape else 1
    kind, addr, nbytes = layout
    if kind == "contiguous":
        if addr == (1 << 64) - 1:  # never written: the fill value, zero by default
            return np.zeros(shape, dtype)
        addr += base
    if nbytes < count * dtype.itemsize:
        raise ValueError("HDF5: dataset storage is smaller than its shape")
    raw = f.d[addr:addr + count * dtype.itemsize]
    return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).copy()
