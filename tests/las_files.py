"""LAS 1.2 / 1.3 / 1.4 files for the ingest tests, packed field by field from the ASPRS layouts (public
header block; point data record formats 0-10, LAS 1.4 R15 tables 7-17) with known integers and
attributes. write_las returns the values it wrote; expected_attributes gives what the reference's
reader (through LASlib's legacy fields) reports for them."""
import numpy as np

FORMAT_SIZE = [20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67]
HEADER_SIZE = {0: 227, 1: 227, 2: 227, 3: 235, 4: 375}


def _put(rec, at, values, dtype):
    b = np.ascontiguousarray(values, dtype=dtype).view(np.uint8).reshape(len(values), np.dtype(dtype).itemsize)
    rec[:, at:at + b.shape[1]] = b


def write_las(path, X, Y, Z, fmt, minor=2, scale=(0.001, 0.001, 0.001), offset=(0.0, 0.0, 0.0), extra=0, vlrs=0,
              seed=0, compressed=False):
    X, Y, Z = (np.asarray(v, np.int64) for v in (X, Y, Z))
    n = len(X)
    rng = np.random.default_rng(seed)
    ext = fmt >= 6
    w = {
        "X": X, "Y": Y, "Z": Z,
        "intensity": rng.integers(0, 65536, n),
        "rn": rng.integers(0, 16 if ext else 8, n), "nr": rng.integers(0, 16 if ext else 8, n),
        "dir": rng.integers(0, 2, n), "edge": rng.integers(0, 2, n),
        "cls": rng.integers(0, 256, n), "ud": rng.integers(0, 256, n), "psid": rng.integers(0, 65536, n),
        "angle": rng.integers(-30000, 30001, n) if ext else rng.integers(-128, 128, n),
        "rgb": rng.integers(0, 65536, (n, 3)), "fmt": fmt,
    }
    rl = FORMAT_SIZE[fmt] + extra
    rec = rng.integers(0, 256, (n, rl)).astype(np.uint8)  # wave packets, GPS, extra bytes: random
    _put(rec, 0, X, "<i4"), _put(rec, 4, Y, "<i4"), _put(rec, 8, Z, "<i4"), _put(rec, 12, w["intensity"], "<u2")
    if not ext:
        flags = w["rn"] | (w["nr"] << 3) | (w["dir"] << 6) | (w["edge"] << 7)
        _put(rec, 14, flags, "<u1"), _put(rec, 15, w["cls"], "<u1"), _put(rec, 16, w["angle"], "<i1")
        _put(rec, 17, w["ud"], "<u1"), _put(rec, 18, w["psid"], "<u2")
        rgb_at = {2: 20, 3: 28, 5: 28}.get(fmt)
    else:
        b15 = rng.integers(0, 64, n) | (w["dir"] << 6) | (w["edge"] << 7)  # class flags, channel, dir, edge
        _put(rec, 14, w["rn"] | (w["nr"] << 4), "<u1"), _put(rec, 15, b15, "<u1"), _put(rec, 16, w["cls"], "<u1")
        _put(rec, 17, w["ud"], "<u1"), _put(rec, 18, w["angle"], "<i2"), _put(rec, 20, w["psid"], "<u2")
        rgb_at = {7: 30, 8: 30, 10: 30}.get(fmt)
    if rgb_at is not None:
        for c in range(3):
            _put(rec, rgb_at + 2 * c, w["rgb"][:, c], "<u2")
    else:
        w["rgb"] = np.zeros((n, 3), np.int64)
    hs = HEADER_SIZE[minor]
    vlr = bytearray()
    for v in range(vlrs):
        payload = bytes(range(10 + v))
        head = bytearray(54)
        head[2:2 + 6] = b"tests\0"
        head[18:20] = (1000 + v).to_bytes(2, "little")
        head[20:22] = len(payload).to_bytes(2, "little")
        vlr += head + payload
    h = bytearray(hs)
    h[0:4] = b"LASF"
    h[24], h[25] = 1, minor
    h[94:96] = hs.to_bytes(2, "little")
    h[96:100] = (hs + len(vlr)).to_bytes(4, "little")
    h[100:104] = vlrs.to_bytes(4, "little")
    h[104] = fmt | (0x80 if compressed else 0)
    h[105:107] = rl.to_bytes(2, "little")
    legacy = n if (fmt < 6 and n < 2 ** 32) else 0
    h[107:111] = legacy.to_bytes(4, "little")
    h[131:155] = np.array(scale, "<f8").tobytes()
    h[155:179] = np.array(offset, "<f8").tobytes()
    if n:
        for a, v in enumerate((X, Y, Z)):
            lo, hi = v.min() * scale[a] + offset[a], v.max() * scale[a] + offset[a]
            h[179 + 16 * a:187 + 16 * a] = np.array([hi], "<f8").tobytes()
            h[187 + 16 * a:195 + 16 * a] = np.array([lo], "<f8").tobytes()
    if minor >= 4:
        h[247:255] = n.to_bytes(8, "little")
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(bytes(vlr))
        f.write(rec.tobytes())
    return w


def expected_attributes(w):
    """(fields n x 8, intensity, psid, rgb) as the reference's readMeta reports them: legacy formats
    carry their fields (5-bit class); LAS 1.4 formats go through LASlib's legacy fields."""
    n = len(w["X"])
    f = np.zeros((n, 8), np.uint8)
    if w["fmt"] < 6:
        f[:, 0], f[:, 1], f[:, 4] = w["rn"], w["nr"], w["cls"] & 31
        f[:, 5] = (w["angle"] & 255)
    else:
        rn, nr = w["rn"], w["nr"]
        f[:, 0] = np.where(nr <= 7, rn & 7, np.where(rn <= 6, rn, np.where(rn >= nr, 7, 6)))
        f[:, 1] = np.minimum(nr, 7)
        f[:, 4] = np.where(w["cls"] <= 31, w["cls"], 0)
        deg = (np.float32(0.006) * w["angle"].astype(np.float32)).astype(np.float32)
        f[:, 5] = np.trunc(deg).astype(np.int64) & 255
    f[:, 2], f[:, 3], f[:, 6] = w["dir"], w["edge"], w["ud"]
    return f, w["intensity"].astype(np.float64), w["psid"].astype(np.uint16), w["rgb"].astype(np.uint32)
