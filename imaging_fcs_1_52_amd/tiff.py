"""16-bit TIFF image stacks (the plugin's input format, read through ImageJ's
getVoxels, IFCS:14797) -> uint16 [T][H][W] for the native pipeline.

Supported: baseline uncompressed grayscale TIFF, 16 bits per sample, one
sample per pixel, little- or big-endian, one page per frame (strips of any
size), and ImageJ's large-stack convention, in which the first page's
description holds "images=N" and the N frames follow the first page's data
contiguously.  When the frames are contiguous in the file the stack is
returned as a read-only memory map (stacks larger than host memory stay
on disk until uploaded).
"""
import re
import struct

import numpy as np

_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}


def _ifds(buf, endian):
    """Yields the tag dictionaries of every page (classic TIFF)."""
    off = struct.unpack(endian + "I", buf[4:8])[0]
    seen = set()
    while off and off not in seen:
        seen.add(off)
        n = struct.unpack(endian + "H", buf[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = off + 2 + 12 * i
            tag, typ, cnt = struct.unpack(endian + "HHI", buf[e:e + 8])
            if typ not in _TYPES:
                continue
            fmt, size = _TYPES[typ]
            nbytes = size * cnt
            data = buf[e + 8:e + 12] if nbytes <= 4 else buf[struct.unpack(endian + "I", buf[e + 8:e + 12])[0]:][:nbytes]
            if typ == 2:
                tags[tag] = bytes(data[:cnt]).rstrip(b"\0").decode("latin-1")
            else:
                tags[tag] = struct.unpack(endian + fmt * cnt, bytes(data[:nbytes]))
        yield tags
        off = struct.unpack(endian + "I", buf[off + 2 + 12 * n:off + 6 + 12 * n])[0]


def read_tiff_stack(path):
    """uint16 array [T][H][W] of a 16-bit grayscale TIFF stack."""
    mm = np.memmap(path, np.uint8, mode="r")
    head = bytes(mm[:4])
    if head[:2] == b"II":
        endian = "<"
    elif head[:2] == b"MM":
        endian = ">"
    else:
        raise ValueError("not a TIFF file")
    if struct.unpack(endian + "H", head[2:4])[0] != 42:
        raise ValueError("only classic (non-BigTIFF) files are supported")
    pages = list(_ifds(mm, endian))
    if not pages:
        raise ValueError("TIFF without pages")
    p0 = pages[0]
    W, H = p0[256][0], p0[257][0]
    if p0.get(258, (1,))[0] != 16 or p0.get(277, (1,))[0] != 1:
        raise ValueError("only 16-bit single-channel stacks are supported")
    if p0.get(259, (1,))[0] != 1:
        raise ValueError("compressed TIFF is not supported")
    dt = np.dtype(endian + "u2")
    frame_bytes = W * H * 2
    desc = p0.get(270, "")
    nimg = re.search(r"images=(\d+)", desc) if isinstance(desc, str) and desc.startswith("ImageJ=") else None
    if nimg and int(nimg.group(1)) > len(pages):
        T = int(nimg.group(1))
        offsets = [p0[273][0] + t * frame_bytes for t in range(T)]
    else:
        T = len(pages)
        offsets = []
        for p in pages:
            if (p[256][0], p[257][0]) != (W, H):
                raise ValueError("pages of different sizes")
            so, sc = p[273], p[279]
            if sum(sc) != frame_bytes:
                raise ValueError("unexpected strip sizes")
            if any(so[i] + sc[i] != so[i + 1] for i in range(len(so) - 1)):
                offsets.append(tuple(zip(so, sc)))   # scattered strips
            else:
                offsets.append(so[0])
    if all(isinstance(o, int) for o in offsets) and \
            all(offsets[t + 1] - offsets[t] == frame_bytes for t in range(T - 1)):
        arr = np.memmap(path, dt, mode="r", offset=offsets[0], shape=(T, H, W))
        return arr if endian == "<" else arr.astype(np.uint16)
    out = np.empty((T, H, W), np.uint16)
    for t, o in enumerate(offsets):
        if isinstance(o, int):
            out[t] = np.frombuffer(mm, dt, count=W * H, offset=o).reshape(H, W)
        else:
            flat = np.concatenate([np.frombuffer(mm, np.uint8, count=c, offset=s) for s, c in o])
            out[t] = flat.view(dt).reshape(H, W)
    return out


def write_tiff_stack(path, stack, imagej=True):
    """Writes uint16 [T][H][W] as an uncompressed little-endian TIFF: one
    page per frame, or (imagej=True) ImageJ's contiguous layout with a
    single page describing all frames."""
    stack = np.ascontiguousarray(stack, np.uint16)
    T, H, W = stack.shape
    desc = ("ImageJ=1.53t\nimages=%d\nslices=%d\nloop=false\n" % (T, T)).encode() + b"\0" if imagej else None
    npages = 1 if imagej else T
    frame_bytes = W * H * 2

    def ifd(data_off, desc_off, next_off):
        entries = [(256, 4, 1, W), (257, 4, 1, H), (258, 3, 1, 16), (259, 3, 1, 1), (262, 3, 1, 1)]
        if desc is not None:
            entries.append((270, 2, len(desc), desc_off))
        entries += [(273, 4, 1, data_off), (277, 3, 1, 1), (278, 4, 1, H), (279, 4, 1, frame_bytes)]
        b = struct.pack("<H", len(entries))
        for tag, typ, cnt, val in entries:
            b += struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", val, 0) if typ == 3 else
                                                      struct.pack("<I", val))
        return b + struct.pack("<I", next_off)

    n_entries = 9 + (1 if desc is not None else 0)
    ifd_size = 2 + 12 * n_entries + 4
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HI", 42, 8))
        meta = 8 + npages * ifd_size
        desc_off = meta
        data0 = meta + (len(desc) if desc is not None else 0)
        data0 += (-data0) % 16
        for p in range(npages):
            nxt = 8 + (p + 1) * ifd_size if p + 1 < npages else 0
            f.write(ifd(data0 + p * frame_bytes, desc_off, nxt))
        if desc is not None:
            f.write(desc)
        f.write(b"\0" * (data0 - f.tell()))
        f.write(stack.tobytes())
