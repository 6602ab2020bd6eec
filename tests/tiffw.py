"""Minimal TIFF writer for tests: uncompressed or caller-encoded strips or
tiles, either byte order, classic TIFF or BigTIFF.  Used for what Pillow's
writer cannot produce (tiles, big-endian samples, BigTIFF) and for EER
containers (compression 65000 / 65001 with raw code streams as strips)."""
import struct


def write(path, dirs, big_endian=False, bigtiff=False):
    """dirs: list of dicts with w, h, bits, fmt (1/2/3), comp, blocks (list of
    bytes: strips top to bottom, or tiles row-major), and either rps (rows per
    strip) or tile (tw, th)."""
    e = ">" if big_endian else "<"
    out = bytearray()
    out += b"MM" if big_endian else b"II"
    if bigtiff:
        out += struct.pack(e + "HHHQ", 43, 8, 0, 0)
        first_ptr = 8
    else:
        out += struct.pack(e + "HI", 42, 0)
        first_ptr = 4
    prev_ptr = first_ptr
    for d in dirs:
        offs = []
        for b in d["blocks"]:
            offs.append(len(out))
            out += b
            if len(out) % 2:
                out += b"\0"
        tags = [(256, 4, [d["w"]]), (257, 4, [d["h"]]), (258, 3, [d["bits"]]), (259, 3, [d["comp"]]),
                (262, 3, [1]), (277, 3, [1]), (339, 3, [d.get("fmt", 1)])]
        if "tile" in d:
            tags += [(322, 4, [d["tile"][0]]), (323, 4, [d["tile"][1]]),
                     (324, 16 if bigtiff else 4, offs), (325, 16 if bigtiff else 4, [len(b) for b in d["blocks"]])]
        else:
            tags += [(273, 16 if bigtiff else 4, offs), (278, 4, [d["rps"]]),
                     (279, 16 if bigtiff else 4, [len(b) for b in d["blocks"]])]
        if "pred" in d:
            tags.append((317, 3, [d["pred"]]))
        tags.sort()
        inl = 8 if bigtiff else 4
        # out-of-line arrays first
        vals = []
        for tag, typ, v in tags:
            fmt = {3: "H", 4: "I", 16: "Q"}[typ]
            raw = struct.pack(e + fmt * len(v), *v)
            if len(raw) > inl:
                if len(out) % 2:
                    out += b"\0"
                ptr = len(out)
                out += raw
                vals.append((tag, typ, len(v), struct.pack(e + ("Q" if bigtiff else "I"), ptr)))
            else:
                vals.append((tag, typ, len(v), raw + b"\0" * (inl - len(raw))))
        if len(out) % 2:
            out += b"\0"
        ifd = len(out)
        if bigtiff:
            out += struct.pack(e + "Q", len(vals))
            for tag, typ, n, raw in vals:
                out += struct.pack(e + "HHQ", tag, typ, n) + raw
            out += struct.pack(e + "Q", 0)
            struct.pack_into(e + "Q", out, prev_ptr, ifd)
            prev_ptr = len(out) - 8
        else:
            out += struct.pack(e + "H", len(vals))
            for tag, typ, n, raw in vals:
                out += struct.pack(e + "HHI", tag, typ, n) + raw
            out += struct.pack(e + "I", 0)
            struct.pack_into(e + "I", out, prev_ptr, ifd)
            prev_ptr = len(out) - 4
    with open(path, "wb") as f:
        f.write(out)
