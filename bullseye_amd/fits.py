"""FITS image export of finalized facet images, without astropy or casacore.

Writes the primary HDU bullseye's exporter produces (bullseye/data/fits_export.py:74-111 through
astropy): data of shape (channels, 1, size_l, size_m) — NAXIS1 = size_m, NAXIS2 = size_l,
NAXIS3 = 1 Stokes axis, NAXIS4 = channels — followed by the same 20 WCS cards in the same order
(RA---SIN / DEC--SIN orthographic axes, STOKES, WAVELENGTH).  Angles are plain floats in degrees
where the reference takes pyrap quantities.  The file layout follows the FITS 3.0 standard (Pence
et al. 2010): 80-character ASCII cards in 2880-byte blocks padded with spaces, big-endian IEEE
data padded with zeros.
"""
import numpy as np

BLOCK = 2880

# FITS 3.0 Table 28 (the reference's FITS_POLARIZATION_CLASSIFIERS)
POLARIZATION_CODES = {"I": 1, "Q": 2, "U": 3, "V": 4, "RR": -1, "LL": -2, "RL": -3, "LR": -4,
                      "XX": -5, "YY": -6, "XY": -7, "YX": -8}


def _value(v):
    """Fixed-format value field (columns 11-30) of a header card."""
    if isinstance(v, (bool, np.bool_)):
        return "%20s" % ("T" if v else "F")
    if isinstance(v, (int, np.integer)):
        return "%20d" % int(v)
    if isinstance(v, (float, np.floating)):
        s = "%.16G" % float(v)
        if "." not in s and "E" not in s and "N" not in s and "I" not in s:
            s += ".0"
        return "%20s" % s
    s = "'%-8s'" % str(v).replace("'", "''")
    return "%-20s" % s


def card(keyword, value=None, comment=None):
    """One 80-character header card."""
    if keyword == "END":
        return "END".ljust(80)
    text = "%-8s= %s" % (keyword, _value(value))
    if comment:
        text += " / " + comment
    if len(text) > 80:
        raise ValueError("header card too long: %r" % text)
    return text.ljust(80)


def header_cards(data, size_l, size_m, cell_l_deg, cell_m_deg, centre_px_l, centre_px_m,
                 pointing_ra_deg, pointing_dec_deg, polarization_term, ref_wavelength,
                 delta_wavelength, channels):
    bitpix = -32 if data.dtype == np.float32 else -64
    cards = [card("SIMPLE", True, "conforms to FITS standard"),
             card("BITPIX", bitpix, "array data type"),
             card("NAXIS", 4, "number of array dimensions"),
             card("NAXIS1", size_m), card("NAXIS2", size_l), card("NAXIS3", 1),
             card("NAXIS4", channels),
             card("EXTEND", True)]
    wcs = [("CRPIX1", centre_px_m, "Pixel coordinate of reference point"),
           ("CDELT1", -float(cell_m_deg), "step per m pixel"),
           ("CTYPE1", "RA---SIN", "Orthog projection"),
           ("CRVAL1", float(pointing_ra_deg), "RA value"),
           ("CUNIT1", "deg", "units are always degrees"),
           ("CRPIX2", centre_px_l, "Pixel coordinate of reference point"),
           ("CDELT2", float(cell_l_deg), "step per l pixel"),
           ("CTYPE2", "DEC--SIN", "Orthog projection"),
           ("CRVAL2", float(pointing_dec_deg), "DEC value"),
           ("CUNIT2", "deg", "units are always degrees"),
           ("CRPIX3", 1, "Pixel coordinate of reference point"),
           ("CDELT3", 1, "dummy value"),
           ("CTYPE3", "STOKES", "Polarization"),
           ("CRVAL3", POLARIZATION_CODES[polarization_term], "Polarization identifier"),
           ("CUNIT3", " ", "Polarization term is unitless"),
           ("CRPIX4", 1, "Pixel coordinate of reference point"),
           ("CDELT4", float(delta_wavelength), "dummy velocity value"),
           ("CTYPE4", "WAVELENGTH", "Wavelength in metres"),
           ("CRVAL4", float(ref_wavelength), "first wavelength in the cube"),
           ("CUNIT4", "m", "wavelength in m")]
    cards += [card(k, v, c) for k, v, c in wcs]
    cards.append(card("END"))
    return cards


def save_to_fits_image(name, size_l, size_m, cell_l_deg, cell_m_deg, centre_px_l, centre_px_m,
                       pointing_ra_deg, pointing_dec_deg, polarization_term, ref_wavelength,
                       delta_wavelength, cube_channel_dim_size, data):
    """fits_export.py:74-111 with angles in degrees; the same argument checks."""
    data = np.asarray(data)
    if data.size != size_l * size_m * cube_channel_dim_size:
        raise Exception("Data size must be equal to size_l * size_m * no_channels_in_cube")
    if not (data.dtype == np.float32 or data.dtype == np.float64):
        raise Exception("Expected float or double typed data but got %s" % data.dtype)
    cube = np.ascontiguousarray(data).reshape(cube_channel_dim_size, 1, size_l, size_m)
    head = "".join(header_cards(cube, size_l, size_m, cell_l_deg, cell_m_deg, centre_px_l,
                                centre_px_m, pointing_ra_deg, pointing_dec_deg,
                                polarization_term, ref_wavelength, delta_wavelength,
                                cube_channel_dim_size)).encode("ascii")
    head += b" " * ((-len(head)) % BLOCK)
    body = cube.astype(cube.dtype.newbyteorder(">")).tobytes()
    body += b"\0" * ((-len(body)) % BLOCK)
    with open(name, "wb") as f:
        f.write(head)
        f.write(body)


def read_fits_image(name):
    """The primary HDU of a file written by save_to_fits_image: (header dict, data array)."""
    raw = open(name, "rb").read()
    hdr, pos = {}, 0
    while True:
        c = raw[pos:pos + 80].decode("ascii")
        pos += 80
        if c.startswith("END"):
            break
        key = c[:8].strip()
        val = c[10:].split(" / ")[0].strip()
        if val.startswith("'"):
            hdr[key] = val[1:val.rindex("'")].rstrip()
        elif val in ("T", "F"):
            hdr[key] = val == "T"
        else:
            hdr[key] = float(val) if any(ch in val for ch in ".EN") else int(val)
    pos = -(-pos // BLOCK) * BLOCK
    dt = np.dtype(">f4") if hdr["BITPIX"] == -32 else np.dtype(">f8")
    shape = tuple(hdr["NAXIS%d" % i] for i in range(hdr["NAXIS"], 0, -1))
    n = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dt, count=n, offset=pos).reshape(shape)
    return hdr, data
