"""Mesh files in the reference's HDF5 format, without an HDF5 library.

``readMesh2dHDF5`` (utilsIO.jl:107-316) reads one dataset per field of
``mesh2d_Int32`` from the root group of ``<name>.h5``: int32 scalars
``nCells, nNodes, nNeibCells, nBSets`` and column-major Julia arrays, which an
HDF5 file stores with the dimensions reversed (a Julia ``(nCells, 4)`` array is
an HDF5 ``(4, nCells)`` dataset with the same bytes).  Three of the file's
names differ from the struct's: ``node2cellL2up``, ``node2cellL2down`` (struct:
``node2cellsL2up/down``).

This module reads and writes the part of the HDF5 format such files use:
superblock version 0 or 1; the root group as a symbol table (version-1
B-tree of symbol-table nodes + local heap, any depth), as link messages in
its header, or as "dense" links in a fractal heap (the layout of the
reference's own mesh file); version-1 object headers with continuation
blocks; datasets with a simple or scalar dataspace,
fixed-point or IEEE floating-point element types of either byte order, and
contiguous or compact storage.  Chunked or filtered datasets, other datatype
classes and version-2 object headers raise ``ValueError``.  The writer emits
exactly that subset (superblock 0, contiguous little-endian datasets), which
any HDF5 library reads.
"""
from __future__ import annotations

import struct

import numpy as np

from .mesh import mesh2d_Int32

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF

# file dataset name -> mesh2d_Int32 field
_FILE_NAMES = {"node2cellL2up": "node2cellsL2up", "node2cellL2down": "node2cellsL2down"}
_SCALARS = ("nCells", "nNodes", "nNeibCells", "nBSets")
_ARRAYS_I32 = ("mesh_connectivity", "bc_data", "bc_indexes", "cell_stiffness", "cell_clusters", "node2cellL2up",
               "node2cellL2down", "cells2nodes")
_ARRAYS_F64 = ("xNodes", "yNodes", "cell_nodes_X", "cell_nodes_Y", "cell_mid_points", "cell_areas", "HX",
               "cell_edges_Nx", "cell_edges_Ny", "cell_edges_length", "node_stencils")


# --------------------------------------------------------------------------- reader
class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if buf[:8] != _SIG:
            raise ValueError("not an HDF5 file (no signature at offset 0)")
        ver = buf[8]
        if ver not in (0, 1):
            raise ValueError(f"HDF5 superblock version {ver} is not supported (0 or 1)")
        self.so, self.sl = buf[13], buf[14]
        if self.so != 8 or self.sl != 8:
            raise ValueError("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)
        self.base = self.u(p, 8)
        root_entry = p + 32
        self.root_ohdr = self.u(root_entry + 8, 8)

    def u(self, off, n):
        return int.from_bytes(self.b[off:off + n], "little")

    def messages(self, addr):
        """(type, data-offset, size) of every message of a version-1 object header."""
        b = self.b
        if b[addr] != 1:
            raise ValueError(f"object header version {b[addr]} at {addr} is not supported (1)")
        nmsg = self.u(addr + 2, 2)
        blocks = [(addr + 16, self.u(addr + 8, 4))]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(out) < nmsg:
                mtype, msize = self.u(p, 2), self.u(p + 2, 2)
                out.append((mtype, p + 8, msize))
                if mtype == 0x10:  # continuation
                    blocks.append((self.u(p + 8, 8), self.u(p + 16, 8)))
                p += 8 + msize
        return out

    def group_entries(self, ohdr):
        msgs = self.messages(ohdr)
        names = {}
        for mtype, off, _ in msgs:
            if mtype == 0x06:  # compact storage: link messages in the header
                name, addr, _ = self._link(off)
                names[name] = addr
        dense = [m for m in msgs if m[0] == 0x02]
        if dense and self.u(dense[0][1] + 2, 8) != _UNDEF:  # link info: dense storage in a fractal heap
            names.update(self._dense_links(dense[0][1]))
        if names:
            return names
        sym = [m for m in msgs if m[0] == 0x11]
        if not sym:
            raise ValueError("root group has neither links nor a symbol table")
        _, off, _ = sym[0]
        btree, heap = self.u(off, 8), self.u(off + 8, 8)
        if self.b[heap:heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        heap_data = self.u(heap + 24, 8)
        names = {}
        self._walk(btree, heap_data, names)
        return names

    def _link(self, off):
        """A link message at ``off``: (name, object header address, length)."""
        b = self.b
        if b[off] != 1:
            raise ValueError(f"link message version {b[off]} is not supported")
        flags = b[off + 1]
        p = off + 2
        ltype = 0
        if flags & 0x08:
            ltype = b[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        nlen_size = 1 << (flags & 3)
        nlen = self.u(p, nlen_size)
        p += nlen_size
        name = b[p:p + nlen].decode()
        p += nlen
        if ltype != 0:
            raise ValueError(f"link '{name}' is not a hard link")
        return name, self.u(p, 8), p + 8 - off

    def _dense_links(self, li):
        """Links kept in a fractal heap (HDF5 1.8 'dense' groups): the heap's
        managed objects are the link messages, packed from the start of each
        direct block in address-space order."""
        b = self.b
        heap = self.u(li + 2, 8)
        if b[heap:heap + 4] != b"FRHP":
            raise ValueError("bad fractal heap")
        filt_len, hflags = self.u(heap + 7, 2), b[heap + 9]
        if filt_len:
            raise ValueError("filtered fractal heaps are not supported")
        nobj = self.u(heap + 14 + 7 * 8, 8)
        p = heap + 14 + 12 * 8
        width, start = self.u(p, 2), self.u(p + 2, 8)
        max_bits = self.u(p + 18, 2)
        root, rows = self.u(p + 22, 8), self.u(p + 30, 2)
        off_size = (max_bits + 7) // 8
        hdr = 5 + 8 + off_size + (4 if hflags & 2 else 0)  # direct block header
        blocks = []
        if rows == 0:
            blocks.append((root, start))
        else:
            if b[root:root + 4] != b"FHIB":
                raise ValueError("bad fractal heap indirect block")
            q = root + 5 + 8 + off_size
            for r in range(rows):
                size = start if r < 2 else start << (r - 1)
                for _ in range(width):
                    blocks.append((self.u(q, 8), size))
                    q += 8
        out = {}
        for addr, size in blocks:
            if addr == _UNDEF or len(out) == nobj:
                continue
            if b[addr:addr + 4] != b"FHDB":
                raise ValueError("bad fractal heap direct block")
            q, end = addr + hdr, addr + size
            while q < end and len(out) < nobj and b[q] == 1:
                name, oh, n = self._link(q)
                out[name] = oh
                q += n
        if len(out) != nobj:
            raise ValueError(f"found {len(out)} of the {nobj} links of a dense group")
        return out

    def _walk(self, node, heap_data, names):
        b = self.b
        if b[node:node + 4] != b"TREE":
            raise ValueError("bad group B-tree node")
        level, used = b[node + 5], self.u(node + 6, 2)
        p = node + 24
        for k in range(used):
            child = self.u(p + 8, 8)
            p += 16
            if level > 0:
                self._walk(child, heap_data, names)
                continue
            if b[child:child + 4] != b"SNOD":
                raise ValueError("bad symbol table node")
            for e in range(self.u(child + 6, 2)):
                q = child + 8 + 40 * e
                noff, oh = self.u(q, 8), self.u(q + 8, 8)
                s = heap_data + noff
                name = b[s:b.index(b"\0", s)].decode()
                names[name] = oh

    def dataset(self, ohdr):
        shape, dt, addr, size, raw = None, None, None, None, None
        for mtype, off, msize in self.messages(ohdr):
            b = self.b
            if mtype == 0x01:
                ver, nd, flags = b[off], b[off + 1], b[off + 2]
                dims_at = off + (8 if ver == 1 else 4)
                if ver == 2 and b[off + 3] == 2:
                    shape = (0,)
                else:
                    shape = tuple(self.u(dims_at + 8 * i, 8) for i in range(nd))
            elif mtype == 0x03:
                cls, bits0, esize = b[off] & 0x0F, b[off + 1], self.u(off + 4, 4)
                order = ">" if bits0 & 1 else "<"
                if cls == 0:
                    dt = np.dtype(f"{order}{'i' if bits0 & 0x08 else 'u'}{esize}")
                elif cls == 1 and esize in (4, 8):
                    dt = np.dtype(f"{order}f{esize}")
                else:
                    raise ValueError(f"datatype class {cls} size {esize} is not supported")
            elif mtype == 0x08:
                ver = b[off]
                if ver in (1, 2):
                    nd, lcls = b[off + 1], b[off + 2]
                    if lcls == 1:
                        addr = self.u(off + 8, 8)
                    elif lcls == 0:
                        n = self.u(off + 8 + 4 * nd, 4)
                        raw = bytes(b[off + 12 + 4 * nd:off + 12 + 4 * nd + n])
                    else:
                        raise ValueError("chunked datasets are not supported")
                elif ver in (3, 4):
                    lcls = b[off + 1]
                    if lcls == 1:
                        addr, size = self.u(off + 2, 8), self.u(off + 10, 8)
                    elif lcls == 0:
                        n = self.u(off + 2, 2)
                        raw = bytes(b[off + 4:off + 4 + n])
                    else:
                        raise ValueError("chunked or virtual datasets are not supported")
                else:
                    raise ValueError(f"layout message version {ver} is not supported")
            elif mtype == 0x0B:
                raise ValueError("filtered (compressed) datasets are not supported")
        if shape is None or dt is None:
            raise ValueError("dataset without dataspace or datatype")
        count = int(np.prod(shape)) if shape else 1
        if raw is not None:
            a = np.frombuffer(raw, dtype=dt, count=count)
        elif addr is None or addr == _UNDEF:
            a = np.zeros(count, dtype=dt)
        else:
            a = np.frombuffer(self.b, dtype=dt, count=count, offset=self.base + addr)
        return a.reshape(shape).astype(dt.newbyteorder("="), copy=True)


def read_hdf5(path: str) -> dict:
    """Every dataset of the root group, as numpy arrays in HDF5 (C-order) shape."""
    with open(path, "rb") as fh:
        r = _Reader(fh.read())
    return {name: r.dataset(oh) for name, oh in r.group_entries(r.root_ohdr).items()}


def read_mesh_hdf5(path: str) -> mesh2d_Int32:
    """``readMesh2dHDF5`` (utilsIO.jl:107-316): a mesh2d_Int32 from ``path``
    (the reference appends ".h5" to its argument; pass the full file name)."""
    d = read_hdf5(path)
    missing = [k for k in _SCALARS + _ARRAYS_I32 + _ARRAYS_F64 if k not in d]
    if missing:
        raise ValueError(f"{path}: datasets missing: {missing}")
    jl = lambda a, t: np.asfortranarray(np.asarray(a, dtype=t).T)  # noqa: E731  (Julia column-major)
    kw = {k: int(np.asarray(d[k]).reshape(-1)[0]) for k in _SCALARS}
    for k in _ARRAYS_I32:
        kw[_FILE_NAMES.get(k, k)] = jl(d[k], np.int32)
    for k in _ARRAYS_F64:
        kw[_FILE_NAMES.get(k, k)] = jl(d[k], np.float64)
    for k in ("xNodes", "yNodes", "cell_areas", "HX", "bc_indexes"):
        kw[k] = np.ascontiguousarray(kw[k]).reshape(-1)
    return mesh2d_Int32(**kw)


# --------------------------------------------------------------------------- writer
def _pad8(n):
    return (n + 7) & ~7


def _msg(mtype, data):
    data = data + b"\0" * (_pad8(len(data)) - len(data))
    return struct.pack("<HHB3x", mtype, len(data), 0) + data


def _ohdr(msgs):
    body = b"".join(msgs)
    return struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4 + body


def _dtype_msg(dt: np.dtype):
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0x00
        return struct.pack("<B3BI", 0x10 | 0, bits, 0, 0, dt.itemsize) + struct.pack("<HH", 0, 8 * dt.itemsize)
    if dt.kind == "f" and dt.itemsize in (4, 8):
        if dt.itemsize == 8:
            sign, eloc, esz, msz, bias = 63, 52, 11, 52, 1023
        else:
            sign, eloc, esz, msz, bias = 31, 23, 8, 23, 127
        return (struct.pack("<B3BI", 0x10 | 1, 0x20, sign, 0, dt.itemsize)
                + struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esz, 0, msz, bias))
    raise ValueError(f"cannot write dtype {dt}")


def write_hdf5(path: str, datasets: dict) -> None:
    """Root-group datasets (name -> numpy array, C-order shape; 0-d for
    scalars) as contiguous little-endian data, superblock version 0."""
    names = sorted(datasets)
    arrays = {}
    for n in names:
        a = np.asarray(datasets[n])
        dt = a.dtype.newbyteorder("<")
        arrays[n] = np.asarray(a, dtype=dt, order="C")  # (ascontiguousarray would make 0-d arrays 1-d)
    # local heap data: "" at 0, then the names, 8-byte aligned
    heap = bytearray(b"\0" * 8)
    name_off = {}
    for n in names:
        name_off[n] = len(heap)
        heap += n.encode() + b"\0"
        heap += b"\0" * (_pad8(len(heap)) - len(heap))
    leaf_k = max(4, (len(names) + 1) // 2)  # one symbol-table node holds 2K entries
    internal_k = 16
    sb_size = 96
    root_ohdr_size = len(_ohdr([_msg(0x11, b"\0" * 16)]))
    heap_hdr = 32
    btree_size = 24 + (2 * internal_k) * 16 + 8
    snod_size = 8 + 2 * leaf_k * 40
    a_root = sb_size
    a_heap = a_root + root_ohdr_size
    a_heapdata = a_heap + heap_hdr
    a_btree = _pad8(a_heapdata + len(heap))
    a_snod = a_btree + btree_size
    p = _pad8(a_snod + snod_size)
    layout = {}
    for n in names:
        a = arrays[n]
        space = struct.pack("<BBBB4x", 1, a.ndim, 0, 0) + b"".join(struct.pack("<Q", s) for s in a.shape)
        fill = struct.pack("<BBBB", 2, 2, 2, 0)
        msgs_wo_layout = [_msg(0x01, space), _msg(0x03, _dtype_msg(a.dtype)), _msg(0x05, fill)]
        hdr_len = len(_ohdr(msgs_wo_layout + [_msg(0x08, b"\0" * 18)]))
        a_hdr = p
        a_data = _pad8(a_hdr + hdr_len)
        layout[n] = (a_hdr, a_data, msgs_wo_layout)
        p = _pad8(a_data + a.nbytes)
    eof = p
    out = bytearray(eof)
    # superblock v0 + root symbol table entry (cache type 1: B-tree and heap)
    out[0:sb_size] = (_SIG + struct.pack("<8B", 0, 0, 0, 0, 0, 8, 8, 0) + struct.pack("<HHI", leaf_k, internal_k, 0)
                      + struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
                      + struct.pack("<QQII", 0, a_root, 1, 0) + struct.pack("<QQ", a_btree, a_heap))
    root = _ohdr([_msg(0x11, struct.pack("<QQ", a_btree, a_heap))])
    out[a_root:a_root + len(root)] = root
    # free-list offset 1 is the format's "no free block" (the heap is exactly full)
    out[a_heap:a_heap + heap_hdr] = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), 1, a_heapdata)
    out[a_heapdata:a_heapdata + len(heap)] = heap
    # one leaf B-tree node with one child: keys "" and the last name
    bt = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, _UNDEF, _UNDEF)
    bt += struct.pack("<QQQ", 0, a_snod, name_off[names[-1]] if names else 0)
    out[a_btree:a_btree + len(bt)] = bt
    sn = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
    for n in names:
        sn += struct.pack("<QQII", name_off[n], layout[n][0], 0, 0) + b"\0" * 16
    out[a_snod:a_snod + len(sn)] = sn
    for n in names:
        a_hdr, a_data, msgs = layout[n]
        a = arrays[n]
        hdr = _ohdr(msgs + [_msg(0x08, struct.pack("<BBQQ", 3, 1, a_data, a.nbytes))])
        out[a_hdr:a_hdr + len(hdr)] = hdr
        out[a_data:a_data + a.nbytes] = a.tobytes()
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def write_mesh_hdf5(path: str, m: mesh2d_Int32) -> None:
    """The inverse of ``read_mesh_hdf5``: the reference's dataset names, Julia
    column-major arrays stored with reversed dimensions, int32 scalars."""
    d = {k: np.int32(getattr(m, k)) for k in _SCALARS}
    for k in _ARRAYS_I32:
        d[k] = np.asarray(getattr(m, _FILE_NAMES.get(k, k)), dtype=np.int32).T
    for k in _ARRAYS_F64:
        d[k] = np.asarray(getattr(m, _FILE_NAMES.get(k, k)), dtype=np.float64).T
    write_hdf5(path, d)


def save_solution_hdf5(path: str, dt: float, flowTime: float, densityCells, UxCells, UyCells,  # noqa: N803
                       pressureCells) -> None:
    """A restart snapshot with the fields of ``solutionCellsT``
    (primeObjects.jl:197-205): dt, flowTime, nCells and the cell primitives.
    (The reference stores that struct with BSON.jl; HDF5 carries the same
    fields in a form any HDF5 reader opens.)"""
    rho = np.asarray(densityCells, dtype=np.float64).reshape(-1)
    write_hdf5(path, {"dt": np.float64(dt), "flowTime": np.float64(flowTime), "nCells": np.int64(rho.size),
                      "densityCells": rho, "UxCells": np.asarray(UxCells, dtype=np.float64).reshape(-1),
                      "UyCells": np.asarray(UyCells, dtype=np.float64).reshape(-1),
                      "pressureCells": np.asarray(pressureCells, dtype=np.float64).reshape(-1)})


def load_solution_hdf5(path: str) -> dict:
    """The snapshot of ``save_solution_hdf5`` (what ``loadPrevResults``,
    loadPrevResults.jl:3-38, takes from it: the cell primitives and flowTime)."""
    d = read_hdf5(path)
    out = {"dt": float(d["dt"]), "flowTime": float(d["flowTime"]), "nCells": int(d["nCells"])}
    for k in ("densityCells", "UxCells", "UyCells", "pressureCells"):
        out[k] = d[k]
        if out[k].size != out["nCells"]:
            raise ValueError(f"{path}: {k} has {out[k].size} values for nCells = {out['nCells']}")
    return out
