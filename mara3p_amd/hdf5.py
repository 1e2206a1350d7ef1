"""A minimal HDF5 writer and reader for checkpoints (no HDF5 library needed).

Writes files in the HDF5 file format (specification version 3.0) using only
the parts a checkpoint needs:

  * superblock version 2 (offsets and lengths of 8 bytes, base address 0);
  * version 2 object headers ("OHDR") with one chunk, Jenkins lookup3
    checksums;
  * groups in compact form: a Link Info message (no fractal heap, no name
    index), a Group Info message and one hard-link message per member;
  * datasets with a version 2 dataspace (scalar or simple), a version 3 fill
    value message and a version 3 contiguous layout;
  * datatypes: IEEE little-endian float64, 32/64-bit signed and unsigned
    integers, fixed-length ASCII strings (null-terminated), one-dimensional
    arrays of a base type (the reference's tuple element, H5Tarray_create of
    n_c doubles) and compounds (the reference's rational iteration count,
    {num: long, den: unsigned long}).

The reader parses exactly that subset (and is what the tests round-trip
through).  Files are built in memory and written at once; a checkpoint of
the 16.8 M-zone minidisk mesh is 403 MB.
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"


# --------------------------------------------------------------------------- checksum
def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 hashlittle, the checksum of HDF5 metadata
    (H5_checksum_lookup3): 32-bit words read little-endian, 12 bytes per mix
    round, the last 1-12 bytes zero-padded into the final round."""
    M = 0xFFFFFFFF

    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & M

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    pos = 0
    while n - pos > 12:
        a = (a + int.from_bytes(data[pos:pos + 4], "little")) & M
        b = (b + int.from_bytes(data[pos + 4:pos + 8], "little")) & M
        c = (c + int.from_bytes(data[pos + 8:pos + 12], "little")) & M
        a = (a - c) & M; a ^= rot(c, 4);  c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 6);  a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 8);  b = (b + a) & M
        a = (a - c) & M; a ^= rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 4);  b = (b + a) & M
        pos += 12
    rest = n - pos
    if rest == 0:
        return c
    tail = bytes(data[pos:]) + bytes(12 - rest)
    a = (a + int.from_bytes(tail[0:4], "little")) & M
    b = (b + int.from_bytes(tail[4:8], "little")) & M
    c = (c + int.from_bytes(tail[8:12], "little")) & M
    c ^= b; c = (c - rot(b, 14)) & M
    a ^= c; a = (a - rot(c, 11)) & M
    b ^= a; b = (b - rot(a, 25)) & M
    c ^= b; c = (c - rot(b, 16)) & M
    a ^= c; a = (a - rot(c, 4)) & M
    b ^= a; b = (b - rot(a, 14)) & M
    c ^= b; c = (c - rot(b, 24)) & M
    return c


# --------------------------------------------------------------------------- datatypes
class Datatype:
    """An HDF5 datatype: `message` is the body of its datatype message,
    `dtype` the numpy element type of one value."""

    def __init__(self, message, dtype):
        self.message = message
        self.dtype = np.dtype(dtype)


def float64():
    # class 1 (floating point) version 1; little-endian, mantissa normalised
    # with the msb implied (bits 4-5 = 2), sign at bit 63; bit offset 0,
    # precision 64, exponent at 52 of 11 bits, mantissa at 0 of 52 bits, bias 1023
    body = bytes([0x11, 0x20, 63, 0]) + struct.pack("<I", 8)
    props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
    return Datatype(body + props, "<f8")


def integer(size, signed):
    # class 0 (fixed point) version 1; little-endian, bit 3 = signed
    body = bytes([0x10, 0x08 if signed else 0x00, 0, 0]) + struct.pack("<I", size)
    props = struct.pack("<HH", 0, 8 * size)
    return Datatype(body + props, ("<i%d" if signed else "<u%d") % size)


def string(length):
    # class 3 (string) version 1; null-terminated padding, ASCII
    length = max(1, length)
    return Datatype(bytes([0x13, 0x00, 0, 0]) + struct.pack("<I", length), "S%d" % length)


def array(base, count):
    # class 10 (array) version 3: rank 1, dimension `count`, then the base type
    size = base.dtype.itemsize * count
    body = bytes([0x3A, 0, 0, 0]) + struct.pack("<I", size) + bytes([1]) + struct.pack("<I", count) + base.message
    return Datatype(body, np.dtype((base.dtype, (count,))))


def compound(members):
    """members: [(name, Datatype)], packed in order without padding."""
    size = sum(t.dtype.itemsize for _, t in members)
    width = 1 if size < 256 else (2 if size < 65536 else 4)
    props = b""
    offset = 0
    fields = []
    for name, t in members:
        props += name.encode("ascii") + b"\x00" + offset.to_bytes(width, "little") + t.message
        fields.append((name, t.dtype))
        offset += t.dtype.itemsize
    n = len(members)
    body = bytes([0x36, n & 0xFF, (n >> 8) & 0xFF, 0]) + struct.pack("<I", size) + props
    return Datatype(body, np.dtype(fields))


# --------------------------------------------------------------------------- writer
def _message(mtype, body, flags=0):
    return struct.pack("<BHB", mtype, len(body), flags) + body


def _object_header(messages):
    chunk = b"".join(messages)
    # version 2, flags: chunk-0 size stored in 4 bytes (bits 0-1 = 2)
    head = b"OHDR" + bytes([2, 0x02]) + struct.pack("<I", len(chunk)) + chunk
    return head + struct.pack("<I", lookup3(head))


class Writer:
    """Builds a file bottom-up: write datasets and subgroups first, then the
    groups that link them, then finish() with the root group's address."""

    def __init__(self):
        self.buf = bytearray(48)            # the superblock, written by finish()

    def _append(self, data, align=8):
        pad = (-len(self.buf)) % align
        self.buf += bytes(pad)
        addr = len(self.buf)
        self.buf += data
        return addr

    def dataset(self, value, dtype=None):
        """A dataset holding `value` (numpy array or scalar) with HDF5 type
        `dtype` (default from the numpy type); returns its header address."""
        if dtype is None:
            dtype = _default_type(value)
        if dtype.dtype.subdtype is not None:
            # array elements: the value's trailing axes are the element's
            base, sub = dtype.dtype.subdtype
            raw = np.array(value, dtype=base, order="C")
            if raw.shape[raw.ndim - len(sub):] != sub:
                raise ValueError("hdf5: the trailing axes %s do not match the array type %s" % (raw.shape, sub))
            shape = raw.shape[:raw.ndim - len(sub)]
        else:
            raw = np.array(value, dtype=dtype.dtype, order="C")   # keeps a scalar 0-d
            shape = raw.shape
        data_addr = self._append(raw.tobytes())
        if shape:
            space = bytes([2, len(shape), 0, 1]) + b"".join(struct.pack("<Q", d) for d in shape)
        else:
            space = bytes([2, 0, 0, 0])      # scalar
        fill = bytes([3, 0x01 | (2 << 2)])  # allocated early, written only if set; no value
        layout = bytes([3, 1]) + struct.pack("<QQ", data_addr, raw.nbytes)
        header = _object_header([_message(0x01, space), _message(0x03, dtype.message, flags=0x01),
                                 _message(0x05, fill, flags=0x01), _message(0x08, layout)])
        return self._append(header)

    def group(self, links):
        """A compact group; links: {name: object header address}."""
        link_info = bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)
        n = len(links)
        max_compact = min(65535, max(8, n + 1))
        group_info = bytes([0, 0x01]) + struct.pack("<HH", max_compact, min(6, max_compact - 1))
        msgs = [_message(0x02, link_info), _message(0x0A, group_info)]
        for name, addr in links.items():
            nb = name.encode("utf-8")
            if not 0 < len(nb) < 256:
                raise ValueError("hdf5: link names must be 1-255 bytes")
            # version 1, flags: bit 4 = character set present (UTF-8)
            msgs.append(_message(0x06, bytes([1, 0x10, 1, len(nb)]) + nb + struct.pack("<Q", addr)))
        return self._append(_object_header(msgs))

    def tree(self, node):
        """Write a nested dict (groups) whose leaves are values or
        (value, Datatype) pairs; returns the address of its group."""
        links = {}
        for name, v in node.items():
            if isinstance(v, dict):
                links[name] = self.tree(v)
            elif isinstance(v, tuple) and len(v) == 2 and isinstance(v[1], Datatype):
                links[name] = self.dataset(v[0], v[1])
            else:
                links[name] = self.dataset(v)
        return self.group(links)

    def finish(self, root_addr):
        sb = SIGNATURE + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", 0, UNDEF, len(self.buf), root_addr)
        sb += struct.pack("<I", lookup3(sb))
        self.buf[:48] = sb
        return bytes(self.buf)


def _default_type(value):
    a = np.asarray(value)
    if a.dtype.kind == "f":
        return float64()
    if a.dtype.kind in "iu":
        return integer(a.dtype.itemsize, a.dtype.kind == "i")
    if a.dtype.kind in "SU":
        return string(len(str(a.item()) if a.dtype.kind == "U" else a.item()))
    raise TypeError("hdf5: no HDF5 type for %r" % a.dtype)


def write_file(path, tree):
    """Write a nested dict as an HDF5 file (see Writer.tree)."""
    w = Writer()
    root = w.tree(tree)
    with open(path, "wb") as f:
        f.write(w.finish(root))


# --------------------------------------------------------------------------- reader
class FormatError(ValueError):
    pass


def _parse_datatype(b, pos):
    """(numpy dtype, bytes consumed) of the datatype message at b[pos:]."""
    cv, f0, f1, f2 = b[pos], b[pos + 1], b[pos + 2], b[pos + 3]
    cls, size = cv & 0x0F, struct.unpack_from("<I", b, pos + 4)[0]
    p = pos + 8
    if cls == 0:
        if f0 & 1:
            raise FormatError("hdf5: big-endian integers are not supported")
        return np.dtype(("<i%d" if f0 & 0x08 else "<u%d") % size), 12
    if cls == 1:
        if size != 8 or (f0 & 1):
            raise FormatError("hdf5: only little-endian float64 is supported")
        return np.dtype("<f8"), 20
    if cls == 3:
        return np.dtype("S%d" % size), 8
    if cls == 10:
        rank = b[p]
        dims = struct.unpack_from("<%dI" % rank, b, p + 1)
        base, used = _parse_datatype(b, p + 1 + 4 * rank)
        return np.dtype((base, dims)), 8 + 1 + 4 * rank + used
    if cls == 6:
        n = f0 | (f1 << 8)
        width = 1 if size < 256 else (2 if size < 65536 else 4)
        fields, offs = [], []
        for _ in range(n):
            end = b.index(b"\x00", p)
            name = b[p:end].decode("ascii")
            p = end + 1
            off = int.from_bytes(b[p:p + width], "little")
            p += width
            t, used = _parse_datatype(b, p)
            p += used
            fields.append((name, t))
            offs.append(off)
        return np.dtype({"names": [f[0] for f in fields], "formats": [f[1] for f in fields],
                         "offsets": offs, "itemsize": size}), p - pos
    raise FormatError("hdf5: datatype class %d is not supported" % cls)


def _object(b, addr):
    """Parse the object header at addr: ("group", {name: addr}) or ("dataset", array)."""
    if b[addr:addr + 4] != b"OHDR" or b[addr + 4] != 2:
        raise FormatError("hdf5: expected a version 2 object header at %d" % addr)
    flags = b[addr + 5]
    p = addr + 6
    if flags & 0x20:
        p += 16
    if flags & 0x10:
        p += 4
    w = 1 << (flags & 0x03)
    size = int.from_bytes(b[p:p + w], "little")
    p += w
    end = p + size
    if struct.unpack_from("<I", b, end)[0] != lookup3(b[addr:end]):
        raise FormatError("hdf5: object header checksum mismatch at %d" % addr)
    links, shape, dtype, layout = {}, None, None, None
    is_group = False
    while p + 4 <= end:
        mtype, msize, _ = struct.unpack_from("<BHB", b, p)
        p += 4 + (2 if flags & 0x04 else 0)
        body = b[p:p + msize]
        if mtype == 0x02:
            is_group = True
            if struct.unpack_from("<Q", body, 2 + (8 if body[1] & 1 else 0))[0] != UNDEF:
                raise FormatError("hdf5: dense link storage is not supported")
        elif mtype == 0x06:
            lf = body[1]
            q = 2
            ltype = 0
            if lf & 0x08:
                ltype = body[q]; q += 1
            if lf & 0x04:
                q += 8
            if lf & 0x10:
                q += 1
            nw = 1 << (lf & 0x03)
            nlen = int.from_bytes(body[q:q + nw], "little")
            q += nw
            name = bytes(body[q:q + nlen]).decode("utf-8")
            q += nlen
            if ltype != 0:
                raise FormatError("hdf5: only hard links are supported")
            links[name] = struct.unpack_from("<Q", body, q)[0]
        elif mtype == 0x01:
            rank, sflags, stype = body[1], body[2], body[3]
            shape = () if stype == 0 else struct.unpack_from("<%dQ" % rank, body, 4)
        elif mtype == 0x03:
            dtype, _ = _parse_datatype(b, p)
        elif mtype == 0x08:
            if body[0] != 3 or body[1] != 1:
                raise FormatError("hdf5: only version 3 contiguous layouts are supported")
            layout = struct.unpack_from("<QQ", body, 2)
        p += msize
    if is_group:
        return "group", links
    if shape is None or dtype is None or layout is None:
        raise FormatError("hdf5: object at %d is neither a group nor a contiguous dataset" % addr)
    count = int(np.prod(shape)) if shape else 1
    a = np.frombuffer(b, dtype=dtype, count=count, offset=layout[0]).reshape(tuple(shape) + dtype.shape)
    return "dataset", a.copy()


def read_file(path):
    """The whole file as a nested dict of numpy arrays (scalars as 0-d arrays)."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != SIGNATURE or b[8] != 2:
        raise FormatError("hdf5: expected a version 2 superblock")
    if struct.unpack_from("<I", b, 44)[0] != lookup3(b[:44]):
        raise FormatError("hdf5: superblock checksum mismatch")
    root = struct.unpack_from("<Q", b, 36)[0]

    def walk(addr):
        kind, v = _object(b, addr)
        if kind == "dataset":
            return v
        return {name: walk(a) for name, a in v.items()}
    return walk(root)
