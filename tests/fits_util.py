"""Writes minimal single-image FITS files (primary HDU only) for the PHOENIX tests:
80-character header cards padded to 2880-byte blocks, big-endian data padded to
2880 bytes, as the FITS standard lays them out."""
import numpy as np

_BITPIX = {np.dtype(">f8"): -64, np.dtype(">f4"): -32, np.dtype(">i4"): 32, np.dtype(">i2"): 16}


def _card(key, value=None, comment=""):
    if value is None:
        s = key.ljust(8)
    else:
        if isinstance(value, bool):
            v = "T" if value else "F"
        elif isinstance(value, str):
            v = "'" + value + "'"
        else:
            v = repr(value) if isinstance(value, float) else str(value)
        s = key.ljust(8) + "= " + v.rjust(20) + (" / " + comment if comment else "")
    return s.ljust(80)[:80]


def write_fits(path, data, extra_cards=0, bscale=None, bzero=None):
    """data: a 1-D or 2-D array in one of the supported dtypes (stored big-endian);
    extra_cards: filler COMMENT cards, to push the header past one block."""
    a = np.asarray(data)
    a = a.astype(a.dtype.newbyteorder(">"))
    cards = [_card("SIMPLE", True), _card("BITPIX", _BITPIX[a.dtype]), _card("NAXIS", a.ndim)]
    for k, n in enumerate(reversed(a.shape)):  # NAXIS1 is the fastest axis
        cards.append(_card(f"NAXIS{k + 1}", int(n)))
    if bscale is not None:
        cards.append(_card("BSCALE", float(bscale)))
    if bzero is not None:
        cards.append(_card("BZERO", float(bzero)))
    for i in range(extra_cards):
        cards.append(("COMMENT filler card %d" % i).ljust(80))
    cards.append(_card("END"))
    head = "".join(cards).encode("ascii")
    head += b" " * (-len(head) % 2880)
    body = a.tobytes()
    body += b"\0" * (-len(body) % 2880)
    with open(path, "wb") as f:
        f.write(head + body)
