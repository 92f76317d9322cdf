"""Helpers for the PCD tests: an LZF encoder (the liblzf format: literal runs of up to 32 bytes, back
references of 3..264 bytes up to 8192 back) and writers of PCD files with arbitrary record layouts."""
import numpy as np


def lzf_encode(data: bytes) -> bytes:
    out = bytearray()
    lit = bytearray()
    table = {}
    i, n = 0, len(data)

    def flush():
        for s in range(0, len(lit), 32):
            chunk = lit[s:s + 32]
            out.append(len(chunk) - 1)
            out.extend(chunk)
        lit.clear()
    while i < n:
        key = data[i:i + 3]
        j = table.get(key) if len(key) == 3 else None
        if len(key) == 3:
            table[key] = i
        if j is not None and i - j - 1 < 8192:
            length = 3
            while i + length < n and length < 264 and data[j + length] == data[i + length]:
                length += 1
            flush()
            dist = i - j - 1
            L = length - 2
            if L < 7:
                out.append((L << 5) | (dist >> 8))
            else:
                out.append((7 << 5) | (dist >> 8))
                out.append(L - 7)
            out.append(dist & 0xff)
            for k in range(i + 1, min(i + length, n - 2)):
                table[data[k:k + 3]] = k
            i += length
        else:
            lit.append(data[i])
            i += 1
    flush()
    return bytes(out)


def write_pcd(path, cols, mode):
    """cols: list of (name, numpy array, pcd type char); mode ascii / binary / binary_compressed."""
    n = len(cols[0][1])
    names = " ".join(c[0] for c in cols)
    sizes = " ".join(str(c[1].dtype.itemsize) for c in cols)
    types = " ".join(c[2] for c in cols)
    head = (f"# .PCD v0.7 - test\nVERSION 0.7\nFIELDS {names}\nSIZE {sizes}\nTYPE {types}\n"
            f"COUNT {' '.join('1' for _ in cols)}\nWIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\n"
            f"DATA {mode}\n").encode()
    with open(path, "wb") as f:
        f.write(head)
        if mode == "ascii":
            for i in range(n):
                f.write((" ".join(repr(float(c[1][i])) for c in cols) + "\n").encode())
        elif mode == "binary":
            rec = np.zeros(n, dtype=[(c[0], c[1].dtype) for c in cols])
            for c in cols:
                rec[c[0]] = c[1]
            f.write(rec.tobytes())
        else:
            raw = b"".join(np.ascontiguousarray(c[1]).tobytes() for c in cols)
            comp = lzf_encode(raw)
            f.write(np.array([len(comp), len(raw)], np.uint32).tobytes())
            f.write(comp)
