"""A reader of the reference's BananaParticleData files, written from the format as the reference
writes it (ParticleSerialization.cpp:612-632 header, 533-577 data; ParticleHelpers.cpp:47-79 for
compressed attributes), independent of the library's own reader: used by the tests to check what
the library writes."""
import numpy as np

_DTYPES = {("real", 4): np.float32, ("real", 8): np.float64, ("uint", 4): np.uint32, ("int", 4): np.int32,
           ("uint16", 2): np.uint16, ("char", 1): np.int8}


def read(path):
    """Returns (header_lines, fixed, particle, n): fixed / particle map names to (type, elem size,
    count, raw bytes) in file order."""
    with open(path, "rb") as f:
        blob = f.read()
    pos = 0
    lines = []
    while True:
        end = blob.index(b"\n", pos)
        line = blob[pos:end].decode("ascii")
        pos = end + 1
        lines.append(line)
        if line == "EndHeader.":
            break
    assert lines[0] == "BananaParticleData"
    fixed, particle, n = {}, {}, None
    order = []
    for line in lines[1:-1]:
        tok = line.split(" ")
        if tok[0] == "NParticles":
            n = int(tok[1])
            continue
        assert tok[0] in ("FixedAttribute", "ParticleAttribute"), line
        name, typ, size, count, nbytes = tok[1], tok[2], int(tok[3]), int(tok[4]), int(tok[5])
        order.append((tok[0] == "FixedAttribute", name, typ, size, count, nbytes))
    # data: fixed attributes, then particle attributes, each group in header (name) order
    for group in (True, False):
        for is_fixed, name, typ, size, count, nbytes in order:
            if is_fixed != group:
                continue
            raw = blob[pos:pos + nbytes]
            assert len(raw) == nbytes, name
            pos += nbytes
            (fixed if is_fixed else particle)[name] = (typ, size, count, raw)
    assert pos == len(blob), "trailing bytes"
    return lines, fixed, particle, n


def values(attr):
    typ, size, count, raw = attr
    return np.frombuffer(raw, _DTYPES[(typ, size)])


def compressed(attr, n):
    """(minima, maxima, uint16 values of shape (n, count)) of a compressed_real attribute."""
    typ, size, count, raw = attr
    assert typ == "compressed_real"
    mn = np.frombuffer(raw[:4 * count], np.float32)
    mx = np.frombuffer(raw[4 * count:8 * count], np.float32)
    q = np.frombuffer(raw[8 * count:], np.uint16).reshape(n, count)
    return mn, mx, q


def compress_reference(v):
    """ParticleHelpers::compress in float arithmetic: (uint16)(65535 * ((v - min) / (max - min)))."""
    v = np.asarray(v, np.float32)
    mn, mx = v.min(0), v.max(0)
    diff = (mx - mn).astype(np.float32)
    t = np.float32(65535.0) * ((v - mn) / diff)
    return mn, mx, t.astype(np.float32).astype(np.uint16)
