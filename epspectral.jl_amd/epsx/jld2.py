"""JLD2 files of scGD0 and ladder results: `save_scgd0` / `load_scgd0` (src/scgd0.jl:375-413) and
`save_scgd0_ladder` / `load_scgd0_ladder` (src/ladder.jl:380-416).

JLD2 files are HDF5 files behind a 512-byte JLD2 header block.  No HDF5 library is used here: this
module writes and reads the subset of the HDF5 format (version 2 superblock, version 2 object
headers with compact link storage, contiguous datasets) that the results need, with the layout of
the reference's writers:

* groups `ωs` and `ωs_spectral` with the keys `ω_origin`, `dω_unit`, `offsets` (the field names of
  `FrequencyGrid`);
* `Σ`, `μ`, `Tlist`, `nlist`, `converged` at the top level, and `history/{err, μ,
  min_spectral_weight, wall, iter_frozen}`;
* for the ladder, `Ωs`, `Tlist`, `μ`, `σ_bubble`, `σ_ladder`, `converged`, `niter`, `dω_unit`,
  `history/norms/{iΩ}_{iT}`, `history/σ_iter/{iΩ}_{iT}` (1-based) and `J` when stored.

Arrays keep Julia's column-major layout: an array of Julia size (n1, n2, ...) is stored with the
HDF5 dimensions reversed (n..., n2, n1), as JLD2 stores it; scalars are scalar dataspaces;
ComplexF64 is the compound {re, im} of two IEEE doubles; Bool is the 8-bit enum {FALSE, TRUE}.
Two things differ from what JLD2 writes: `el_i` (an ElectronPhonon `BandStates`; the states here
belong to the setup, not to the result) is not written, and each `σ_iter` entry (a
`Vector{Matrix}` in the reference, stored by JLD2 as object references) is one dense
[3, 3, niter] array.
"""
import struct

import numpy as np

from .grids import FrequencyGrid

JLD2_HEADER = b"HDF5-based Julia Data Format, version 0.1.1"
HEADER_LEN = 512
HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


# ---- Jenkins lookup3 (hashlittle), the checksum of HDF5's version-2 metadata -----------------

def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def lookup3(data: bytes, initval: int = 0) -> int:
    """Bob Jenkins' lookup3 `hashlittle` of `data` (the HDF5 metadata checksum, initval 0)."""
    M = 0xFFFFFFFF
    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    off = 0
    while n > 12:
        a = (a + int.from_bytes(data[off:off + 4], "little")) & M
        b = (b + int.from_bytes(data[off + 4:off + 8], "little")) & M
        c = (c + int.from_bytes(data[off + 8:off + 12], "little")) & M
        a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M
        off += 12
        n -= 12
    if n == 0:
        return c
    tail = data[off:] + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & M
    b = (b + int.from_bytes(tail[4:8], "little")) & M
    c = (c + int.from_bytes(tail[8:12], "little")) & M
    c ^= b; c = (c - _rot(b, 14)) & M
    a ^= c; a = (a - _rot(c, 11)) & M
    b ^= a; b = (b - _rot(a, 25)) & M
    c ^= b; c = (c - _rot(b, 16)) & M
    a ^= c; a = (a - _rot(c, 4)) & M
    b ^= a; b = (b - _rot(a, 14)) & M
    c ^= b; c = (c - _rot(b, 24)) & M
    return c


# ---- datatypes --------------------------------------------------------------------------------

def _dt_float64():
    # class 1 (floating point), version 1; little endian, implied msb, sign at bit 63
    return (bytes([0x11, 0x20, 63, 0]) + struct.pack("<I", 8) +
            struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023))


def _dt_int(size, signed):
    return bytes([0x10, 0x08 if signed else 0x00, 0, 0]) + struct.pack("<I", size) + struct.pack("<HH", 0, 8 * size)


def _dt_complex128():
    # class 6 (compound), version 3: members re (offset 0) and im (offset 8), 1-byte offsets
    body = b"re\x00" + bytes([0]) + _dt_float64() + b"im\x00" + bytes([8]) + _dt_float64()
    return bytes([0x36, 2, 0, 0]) + struct.pack("<I", 16) + body


def _dt_bool():
    # class 8 (enumeration), version 3: base uint8, FALSE = 0, TRUE = 1
    body = _dt_int(1, False) + b"FALSE\x00TRUE\x00" + bytes([0, 1])
    return bytes([0x38, 2, 0, 0]) + struct.pack("<I", 1) + body


def _datatype_for(a):
    if a.dtype == np.float64:
        return _dt_float64()
    if a.dtype == np.complex128:
        return _dt_complex128()
    if a.dtype == np.bool_:
        return _dt_bool()
    if a.dtype.kind in "iu":
        return _dt_int(a.dtype.itemsize, a.dtype.kind == "i")
    raise TypeError(f"no HDF5 datatype for {a.dtype}")


def _parse_datatype(buf, p):
    """(numpy dtype, bytes consumed) of the datatype message at buf[p:]."""
    cls, ver = buf[p] & 0x0F, buf[p] >> 4
    bits = buf[p + 1] | (buf[p + 2] << 8) | (buf[p + 3] << 16)
    size = struct.unpack_from("<I", buf, p + 4)[0]
    q = p + 8
    if cls == 0:
        signed = bool(bits & 0x08)
        return np.dtype(("<i" if signed else "<u") + str(size)), 8 + 4
    if cls == 1:
        if size != 8:
            raise ValueError("only IEEE double floating-point datasets are read")
        return np.dtype("<f8"), 8 + 12
    if cls == 6:
        nmem = bits & 0xFFFF
        if ver != 3:
            raise ValueError("only version-3 compound datatypes are read")
        names = []
        for _ in range(nmem):
            e = buf.index(b"\x00", q)
            names.append(buf[q:e].decode())
            q = e + 1
            q += 1 if size < 256 else (2 if size < 65536 else 4)
            _, used = _parse_datatype(buf, q)
            q += used
        if names != ["re", "im"] or size != 16:
            raise ValueError(f"unknown compound {names}")
        return np.dtype("<c16"), q - p
    if cls == 8:
        nmem = bits & 0xFFFF
        base, used = _parse_datatype(buf, q)
        q += used
        for _ in range(nmem):
            q = buf.index(b"\x00", q) + 1
        q += nmem * base.itemsize
        return np.dtype(np.bool_), q - p
    raise ValueError(f"datatype class {cls} is not read")


# ---- the writer --------------------------------------------------------------------------------

def _message(mtype, data):
    return struct.pack("<BHB", mtype, len(data), 0) + data


def _object_header(messages):
    body = b"".join(messages)
    head = b"OHDR" + bytes([2, 2]) + struct.pack("<I", len(body))
    blob = head + body
    return blob + struct.pack("<I", lookup3(blob))


class _Writer:
    """Objects appended after the superblock; addresses are relative to the superblock (the base
    address), which follows the 512-byte header."""

    def __init__(self):
        self.buf = bytearray()
        self.base = HEADER_LEN + 48          # header block, then the 48-byte superblock

    def here(self):
        return 48 + len(self.buf)            # relative to the superblock

    def append(self, b):
        addr = self.here()
        self.buf += b
        return addr

    def dataset(self, value):
        a = np.asarray(value)
        if a.dtype == np.float32:
            a = a.astype(np.float64)
        if a.ndim == 0:
            space = bytes([2, 0, 0, 0])
            raw = a.reshape(1).tobytes()
        else:
            space = bytes([2, a.ndim, 0, 1]) + b"".join(struct.pack("<Q", d) for d in reversed(a.shape))
            raw = np.asarray(a).ravel(order="F").tobytes()
        if a.dtype == np.bool_:
            raw = a.astype(np.uint8).ravel(order="F").tobytes() if a.ndim else bytes([int(a)])
        data_addr = self.append(raw) if raw else UNDEF
        msgs = [_message(0x01, space), _message(0x03, _datatype_for(a)),
                _message(0x05, bytes([3, 0x05])),
                _message(0x08, bytes([3, 1]) + struct.pack("<QQ", data_addr, len(raw)))]
        return self.append(_object_header(msgs))

    def group(self, tree):
        links = []
        for name, value in tree.items():
            addr = self.group(value) if isinstance(value, dict) else self.dataset(value)
            nb = name.encode("utf-8")
            if len(nb) > 255:
                raise ValueError("link names are at most 255 bytes")
            # version 1, flags: 1-byte name length, character set present (UTF-8)
            links.append(_message(0x06, bytes([1, 0x10, 1, len(nb)]) + nb + struct.pack("<Q", addr)))
        msgs = [_message(0x02, bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)),
                _message(0x0A, bytes([0, 0]))] + links
        return self.append(_object_header(msgs))


def write_tree(filename, tree: dict):
    """Write a nested dict of arrays / scalars (dicts are groups) as a JLD2-headed HDF5 file."""
    w = _Writer()
    root = w.group(tree)
    eof = w.here()
    sb = HDF5_SIGNATURE + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", HEADER_LEN, UNDEF, eof, root)
    sb += struct.pack("<I", lookup3(sb))
    head = JLD2_HEADER + bytes(HEADER_LEN - len(JLD2_HEADER))
    with open(filename, "wb") as f:
        f.write(head + sb + bytes(w.buf))
    return filename


# ---- the reader --------------------------------------------------------------------------------

def _read_object(buf, base, addr):
    p = base + addr
    if buf[p:p + 4] != b"OHDR" or buf[p + 4] != 2:
        raise ValueError(f"no version-2 object header at {addr}")
    flags = buf[p + 5]
    q = p + 6
    if flags & 0x20:
        q += 16
    if flags & 0x10:
        q += 4
    w = 1 << (flags & 3)
    size = int.from_bytes(buf[q:q + w], "little")
    q += w
    end = q + size
    if lookup3(bytes(buf[p:end])) != struct.unpack_from("<I", buf, end)[0]:
        raise ValueError(f"object header checksum mismatch at {addr}")
    msgs = []
    while q + 4 <= end:
        mtype, msize, mflags = struct.unpack_from("<BHB", buf, q)
        q += 4 + (2 if flags & 0x04 else 0)
        msgs.append((mtype, q))
        q += msize
    links = {}
    space = dtype = layout = None
    for mtype, m in msgs:
        if mtype == 0x06:
            lf = buf[m + 1]
            r = m + 2
            ltype = 0
            if lf & 0x08:
                ltype = buf[r]; r += 1
            if lf & 0x04:
                r += 8
            if lf & 0x10:
                r += 1
            nl = 1 << (lf & 3)
            n = int.from_bytes(buf[r:r + nl], "little")
            r += nl
            name = bytes(buf[r:r + n]).decode("utf-8")
            r += n
            if ltype != 0:
                raise ValueError("only hard links are read")
            links[name] = struct.unpack_from("<Q", buf, r)[0]
        elif mtype == 0x01:
            rank, dflags, stype = buf[m + 1], buf[m + 2], buf[m + 3]
            dims = [struct.unpack_from("<Q", buf, m + 4 + 8 * k)[0] for k in range(rank)]
            space = (stype, dims)
        elif mtype == 0x03:
            dtype, _ = _parse_datatype(buf, m)
        elif mtype == 0x08:
            if buf[m] != 3 or buf[m + 1] != 1:
                raise ValueError("only contiguous version-3 layouts are read")
            layout = struct.unpack_from("<QQ", buf, m + 2)
    if space is None:
        return {name: _read_object(buf, base, a) for name, a in links.items()}
    stype, dims = space
    addr_d, nbytes = layout
    count = int(np.prod(dims)) if stype == 1 else 1
    itemsize = 1 if dtype == np.bool_ else dtype.itemsize
    raw = bytes(buf[base + addr_d: base + addr_d + nbytes]) if nbytes else b""
    if len(raw) != count * itemsize:
        raise ValueError("dataset size mismatch")
    arr = np.frombuffer(raw, dtype=np.uint8 if dtype == np.bool_ else dtype).copy()
    if dtype == np.bool_:
        arr = arr.astype(np.bool_)
    if stype == 0:
        return arr[0]
    return arr.reshape(tuple(reversed(dims)), order="F")


def read_tree(filename):
    """Read a file written by `write_tree` back into a nested dict."""
    with open(filename, "rb") as f:
        buf = f.read()
    p = buf.find(HDF5_SIGNATURE)
    if p < 0 or buf[p + 8] not in (2, 3):
        raise ValueError(f"{filename}: no version-2 HDF5 superblock")
    sb = buf[p:p + 48]
    if lookup3(sb[:44]) != struct.unpack_from("<I", sb, 44)[0]:
        raise ValueError(f"{filename}: superblock checksum mismatch")
    base, _, eof, root = struct.unpack_from("<QQQQ", sb, 12)
    return _read_object(buf, base, root)


# ---- the result writers and readers -------------------------------------------------------------

def _grid_group(g: FrequencyGrid):
    return {"ω_origin": np.float64(g.ω_origin), "dω_unit": np.float64(g.dω_unit),
            "offsets": np.asarray(g.offsets, dtype=np.int64)}


def _grid_from(d):
    return FrequencyGrid(float(d["ω_origin"]), float(d["dω_unit"]), np.asarray(d["offsets"], dtype=np.int64))


def save_scgd0(filename, res):
    """`save_scgd0(filename, res)` (src/scgd0.jl:375-388): the grids as groups keyed by the field
    names of FrequencyGrid, the arrays, `history` as a group.  Returns filename."""
    tree = {"ωs": _grid_group(res.ωs), "ωs_spectral": _grid_group(res.ωs_spectral),
            "Σ": np.asarray(res.Σ, dtype=np.complex128), "μ": np.asarray(res.μ, dtype=np.float64),
            "Tlist": np.asarray(res.Tlist, dtype=np.float64), "nlist": np.asarray(res.nlist, dtype=np.float64),
            "converged": np.asarray(res.converged, dtype=np.bool_),
            "history": {k: np.asarray(res.history[k]) for k in
                        ("err", "μ", "min_spectral_weight", "wall", "iter_frozen")}}
    return write_tree(filename, tree)


def load_scgd0(filename):
    """`load_scgd0(filename)` (src/scgd0.jl:395-404)."""
    from .scgd0 import SCGD0Result
    d = read_tree(filename)
    return SCGD0Result(_grid_from(d["ωs"]), _grid_from(d["ωs_spectral"]), d["Σ"], d["μ"], d["Tlist"], d["nlist"],
                       d["converged"], dict(d["history"]))


def save_scgd0_ladder(filename, res):
    """`save_scgd0_ladder(filename, res)` (src/ladder.jl:386-398)."""
    nO, nT = np.asarray(res.converged).shape
    norms, sit = {}, {}
    for iO in range(nO):
        for iT in range(nT):
            h = res.history[iO][iT]
            norms[f"{iO + 1}_{iT + 1}"] = np.asarray(h["norms"], dtype=np.float64)
            s = [np.asarray(m, dtype=np.float64) for m in h["σ_iter"]]
            sit[f"{iO + 1}_{iT + 1}"] = (np.stack(s, axis=2) if s else np.zeros((3, 3, 0)))
    tree = {"Ωs": np.asarray(res.Ωs, dtype=np.float64), "Tlist": np.asarray(res.Tlist, dtype=np.float64),
            "μ": np.asarray(res.μ, dtype=np.float64), "σ_bubble": np.asarray(res.σ_bubble, dtype=np.float64),
            "σ_ladder": np.asarray(res.σ_ladder, dtype=np.float64),
            "converged": np.asarray(res.converged, dtype=np.bool_), "niter": np.asarray(res.niter, dtype=np.int64),
            "dω_unit": np.float64(res.dω_unit), "history": {"norms": norms, "σ_iter": sit}}
    if res.J is not None:
        tree["J"] = np.asarray(res.J, dtype=np.complex128)
    return write_tree(filename, tree)


def load_scgd0_ladder(filename):
    """`load_scgd0_ladder(filename)` (src/ladder.jl:405-416)."""
    from .ladder import SCGD0LadderResult
    d = read_tree(filename)
    nO, nT = len(d["Ωs"]), len(d["Tlist"])
    h = d["history"]
    history = [[dict(norms=h["norms"][f"{iO + 1}_{iT + 1}"],
                     σ_iter=[h["σ_iter"][f"{iO + 1}_{iT + 1}"][:, :, k]
                             for k in range(h["σ_iter"][f"{iO + 1}_{iT + 1}"].shape[2])])
                for iT in range(nT)] for iO in range(nO)]
    return SCGD0LadderResult(d["Ωs"], d["Tlist"], d["μ"], d["σ_bubble"], d["σ_ladder"], d["converged"], d["niter"],
                             history, float(d["dω_unit"]), d.get("J"))
