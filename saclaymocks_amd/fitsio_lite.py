"""Minimal FITS codec for the files on the hot path (fitsio/astropy are absent).

Covers what the reference reads and writes around the box and forest stages:
  - images (BITPIX 8/16/32/64/-32/-64), e.g. box-<ix>.fits (make_boxes.py:110-117)
  - binary tables with scalar or fixed-repeat columns of types L, B, I, J, K,
    E, D and A (strings), e.g. etc/params.fits, the METADATA tables of
    spectra-*.fits.gz (make_spectra.py:566-597) and spectra_merged-*.fits.gz
    (merge_spectra.py:385-406)
  - gzip-compressed files (name ending in .gz)
Data are big-endian on disk; headers are 80-character cards in 2880-byte blocks.
"""
import gzip
import io

import numpy as np

BLOCK = 2880
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_BITPIX_OF = {np.dtype(v).newbyteorder("="): k for k, v in _BITPIX.items()}
_TFORM = {"L": ("i1", 1), "B": ("u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8), "E": (">f4", 4),
          "D": (">f8", 8), "A": ("S", 1)}


def _open(path, mode):
    if str(path).endswith(".gz"):
        return gzip.open(path, mode)
    return open(path, mode)


# ------------------------------------------------------------------ reading
def _parse_value(v):
    v = v.strip()
    if not v:
        return None
    if v.startswith("'"):
        end = v.find("'", 1)
        while end != -1 and end + 1 < len(v) and v[end + 1] == "'":
            end = v.find("'", end + 2)
        return v[1:end].replace("''", "'").rstrip()
    if "/" in v:
        v = v[:v.index("/")].strip()
        if not v:
            return None                    # undefined value
    if v == "T":
        return True
    if v == "F":
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E"))
    except ValueError:
        return v


class Header(dict):
    """Header cards by keyword; lookups ignore case, as FITS keywords do."""

    def __getitem__(self, k):
        return dict.__getitem__(self, k.upper())

    def get(self, k, default=None):
        return dict.get(self, k.upper(), default)

    def __contains__(self, k):
        return dict.__contains__(self, k.upper())


def _read_header(f):
    """Parse one header from the file object positioned at its first block;
    None at the end of the file."""
    hdr = Header()
    first = True
    while True:
        blk = f.read(BLOCK)
        if first and (len(blk) < 8 or not blk[:8].strip(b" \0")):
            return None
        first = False
        if len(blk) < BLOCK:
            raise ValueError("truncated FITS header")
        for i in range(0, BLOCK, 80):
            card = blk[i:i + 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                return hdr
            if card[8:10] == "= ":
                hdr[key] = _parse_value(card[10:])


def _data_bytes(h):
    naxis = h.get("NAXIS", 0)
    if naxis == 0:
        return 0
    n = 1
    for i in range(1, naxis + 1):
        n *= h["NAXIS%d" % i]
    return abs(h["BITPIX"]) // 8 * h.get("GCOUNT", 1) * (h.get("PCOUNT", 0) + n)


def _table_dtype(h):
    fields = []
    for i in range(1, h["TFIELDS"] + 1):
        tf = str(h["TFORM%d" % i]).strip()
        j = 0
        while j < len(tf) and tf[j].isdigit():
            j += 1
        rep = int(tf[:j]) if j else 1
        code = tf[j]
        if code not in _TFORM:
            raise ValueError(f"unsupported TFORM {tf}")
        base, _ = _TFORM[code]
        name = str(h.get("TTYPE%d" % i, "col%d" % i)).strip()
        if code == "A":
            fields.append((name, "S%d" % rep))
        elif rep == 1:
            fields.append((name, base))
        else:
            fields.append((name, base, (rep,)))
    return np.dtype(fields)


def _logical_columns(h):
    """Names of the TFORM 'L' columns: one byte per value, 'T' true, 'F'
    false, 0 null (FITS standard 4.0, section 7.3.3.1)."""
    out = []
    for i in range(1, h["TFIELDS"] + 1):
        if str(h["TFORM%d" % i]).strip().lstrip("0123456789").startswith("L"):
            out.append(str(h.get("TTYPE%d" % i, "col%d" % i)).strip())
    return out


def _decode(h, raw):
    if h.get("XTENSION", "").strip() == "BINTABLE":
        rec = np.frombuffer(raw, dtype=_table_dtype(h), count=h["NAXIS2"]).copy()
        logical = _logical_columns(h)
        if logical:
            dt = np.dtype([(n, ("?", rec.dtype[n].shape) if n in logical else rec.dtype[n])
                           for n in rec.dtype.names])
            out = np.zeros(len(rec), dtype=dt)
            for n in rec.dtype.names:
                out[n] = (rec[n] == ord("T")) if n in logical else rec[n]
            rec = out
        return rec
    if h.get("NAXIS", 0) > 0:
        shape = tuple(h["NAXIS%d" % i] for i in range(h["NAXIS"], 0, -1))
        data = np.frombuffer(raw, dtype=_BITPIX[h["BITPIX"]]).reshape(shape)
        bscale, bzero = h.get("BSCALE", 1), h.get("BZERO", 0)
        data = data.astype(data.dtype.newbyteorder("="))
        if bscale != 1 or bzero != 0:
            data = data * bscale + bzero
        return data
    return None


class HDU:
    def __init__(self, header, data):
        self.header = header
        self.data = data

    @property
    def name(self):
        return str(self.header.get("EXTNAME", "")).strip()


def _scan(path, want):
    """Walk the HDUs of a file; decode the data of those for which
    want(index, header) is true and skip the others' data blocks."""
    out = []
    with _open(path, "rb") as f:
        i = 0
        while True:
            h = _read_header(f)
            if h is None:
                break
            nbytes = _data_bytes(h)
            padded = (nbytes + BLOCK - 1) // BLOCK * BLOCK
            if want(i, h):
                raw = f.read(nbytes)
                if len(raw) < nbytes:
                    raise ValueError("truncated FITS data")
                f.seek(padded - nbytes, 1)
                out.append(HDU(h, _decode(h, raw)))
            else:
                f.seek(padded, 1)
                out.append(HDU(h, None))
            i += 1
    return out


def read_all(path):
    """Every HDU of a FITS file: list of HDU(header dict, numpy data or None)."""
    return _scan(path, lambda i, h: True)


def headers(path):
    """Headers of every HDU, without reading their data."""
    return [h.header for h in _scan(path, lambda i, h: False)]


def _match(ext, i, h):
    if isinstance(ext, str):
        return str(h.get("EXTNAME", "")).strip().upper() == ext.upper()
    return i == ext


def read(path, ext=None):
    """fitsio.read: the data of `ext` (index or EXTNAME); default the first HDU
    with data.  Only that HDU's data is read from the file."""
    if ext is None:
        for i, h in enumerate(headers(path)):
            if _data_bytes(h) > 0:
                ext = i
                break
        else:
            return None
    hdus = _scan(path, lambda i, h: _match(ext, i, h))
    for i, h in enumerate(hdus):
        if _match(ext, i, h.header):
            return h.data
    raise KeyError(f"no HDU {ext!r} in {path}")


def read_header(path, ext=0):
    for i, h in enumerate(headers(path)):
        if _match(ext, i, h):
            return h
    raise KeyError(f"no HDU {ext!r} in {path}")


# ------------------------------------------------------------------ writing
def _card(key, value, comment=None):
    if value is None:                      # keyword with an undefined value
        card = "%-8s= %20s" % (key[:8], "")
        if comment:
            card += " / " + comment
        return card[:80].ljust(80)
    if isinstance(value, (bool, np.bool_)):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % int(value)
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper().replace("INF", "Inf")
    else:
        s = str(value).replace("'", "''")
        v = "'%-8s'" % s
    card = "%-8s= %s" % (key[:8], v)
    if comment:
        card += " / " + comment
    return card[:80].ljust(80)


def _header_bytes(cards):
    text = "".join(cards) + "END".ljust(80)
    pad = (-len(text)) % BLOCK
    return (text + " " * pad).encode("ascii")


def _pad(b, fill=b"\0"):
    return b + fill * ((-len(b)) % BLOCK)


def _user_cards(header):
    cards = []
    if header is None:
        return cards
    items = header.items() if isinstance(header, dict) else [(d["name"], d["value"]) if "comment" not in d
                                                             else (d["name"], (d["value"], d["comment"]))
                                                             for d in header]
    for k, v in items:
        if isinstance(v, tuple):
            cards.append(_card(k.upper(), v[0], v[1]))
        else:
            cards.append(_card(k.upper(), v))
    return cards


def _image_hdu(arr, header, extname, primary):
    arr = np.asarray(arr)
    dt = arr.dtype.newbyteorder("=")
    if dt == np.dtype("bool"):
        arr = arr.astype(np.uint8)
        dt = np.dtype("u1")
    bitpix = _BITPIX_OF[dt]
    cards = [_card("SIMPLE", True) if primary else _card("XTENSION", "IMAGE"), _card("BITPIX", bitpix),
             _card("NAXIS", arr.ndim)]
    for i, n in enumerate(arr.shape[::-1]):
        cards.append(_card("NAXIS%d" % (i + 1), n))
    if primary:
        cards.append(_card("EXTEND", True))
    else:
        cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
    if extname:
        cards.append(_card("EXTNAME", extname))
    cards += _user_cards(header)
    data = np.ascontiguousarray(arr, dtype=_BITPIX[bitpix]).tobytes()
    return _header_bytes(cards) + _pad(data)


def _tform(dt):
    base = dt.base
    rep = int(np.prod(dt.shape)) if dt.shape else 1
    k = base.kind
    if k == "S":
        return "%dA" % base.itemsize, "S%d" % base.itemsize
    if k == "U":
        return "%dA" % (base.itemsize // 4), None
    m = {("f", 8): "D", ("f", 4): "E", ("i", 8): "K", ("i", 4): "J", ("i", 2): "I", ("u", 1): "B",
         ("b", 1): "L", ("i", 1): "B"}
    code = m.get((k, base.itemsize))
    if code is None:
        raise ValueError(f"unsupported column dtype {dt}")
    return ("%d%s" % (rep, code) if rep > 1 else code), _TFORM[code][0]


def _table_hdu(rec, header, extname):
    rec = np.asarray(rec)
    names = rec.dtype.names
    fields = []
    cards_cols = []
    for i, n in enumerate(names):
        dt = rec.dtype[n]
        if dt.kind == "U":
            col = np.char.encode(rec[n], "ascii")
            dt = col.dtype
        tf, base = _tform(dt)
        cards_cols += [_card("TTYPE%d" % (i + 1), n), _card("TFORM%d" % (i + 1), tf)]
        if dt.base.kind == "S":
            fields.append((n, dt.base.str))
        elif dt.shape:
            fields.append((n, base, dt.shape))
        else:
            fields.append((n, base))
    out = np.zeros(len(rec), dtype=np.dtype(fields))
    for n in names:
        col = rec[n]
        if col.dtype.kind == "U":
            col = np.char.encode(col, "ascii")
        elif col.dtype.kind == "b":
            col = np.where(col, ord("T"), ord("F")).astype(np.int8)
        out[n] = col
    cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", out.dtype.itemsize), _card("NAXIS2", len(out)), _card("PCOUNT", 0),
             _card("GCOUNT", 1), _card("TFIELDS", len(names))] + cards_cols
    if extname:
        cards.append(_card("EXTNAME", extname))
    cards += _user_cards(header)
    return _header_bytes(cards) + _pad(out.tobytes())


class FITS:
    """Write-only counterpart of fitsio.FITS(path, 'rw', clobber=True)."""

    def __init__(self, path, mode="rw", clobber=True):
        self.path = path
        self._buf = io.BytesIO()
        self._n = 0

    def write(self, data, header=None, extname=None, names=None):
        """Image (ndarray), table (structured array), or a list of columns with `names`."""
        if names is not None:
            cols = [np.asarray(c) for c in data]
            rec = np.zeros(len(cols[0]), dtype=[(n, c.dtype, c.shape[1:]) for n, c in zip(names, cols)])
            for n, c in zip(names, cols):
                rec[n] = c
            data = rec
        data = np.asarray(data)
        if self._n == 0 and data.dtype.names is not None:
            self._buf.write(_header_bytes([_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0),
                                           _card("EXTEND", True)]))
            self._n = 1
        if data.dtype.names is not None:
            self._buf.write(_table_hdu(data, header, extname))
        else:
            self._buf.write(_image_hdu(data, header, extname, primary=(self._n == 0)))
        self._n += 1

    def close(self):
        with _open(self.path, "wb") as f:
            f.write(self._buf.getvalue())

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
