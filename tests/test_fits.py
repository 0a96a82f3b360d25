"""FITS export (bullseye_amd.fits) against the FITS 3.0 layout rules and the reference's cards
(bullseye/data/fits_export.py:74-111); astropy is absent, so the file is checked structurally
and read back with the module's reader."""
import os

import numpy as np
import pytest

from bullseye_amd import fits


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_round_trip_and_layout(tmp_path, dtype):
    rng = np.random.default_rng(1)
    L, M, C = 24, 20, 3
    img = rng.normal(size=(C, L, M)).astype(dtype)
    name = os.path.join(tmp_path, "img.fits")
    fits.save_to_fits_image(name, L, M, 2 / 3600.0, 2 / 3600.0, L // 2, M // 2, 45.0, -30.0, "I",
                            0.21, 1e-4, C, img)
    raw = open(name, "rb").read()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
    assert cards[0] == "SIMPLE  =                    T / conforms to FITS standard".ljust(80)
    assert cards[1].startswith("BITPIX  = %20d" % (-32 if dtype == np.float32 else -64))
    keys = [c[:8].strip() for c in cards if c.strip()]
    assert keys[:8] == ["SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "NAXIS3", "NAXIS4",
                        "EXTEND"]
    assert keys[8:28] == ["CRPIX1", "CDELT1", "CTYPE1", "CRVAL1", "CUNIT1", "CRPIX2", "CDELT2",
                          "CTYPE2", "CRVAL2", "CUNIT2", "CRPIX3", "CDELT3", "CTYPE3", "CRVAL3",
                          "CUNIT3", "CRPIX4", "CDELT4", "CTYPE4", "CRVAL4", "CUNIT4"]
    assert keys[28] == "END"
    hdr, data = fits.read_fits_image(name)
    assert (hdr["NAXIS1"], hdr["NAXIS2"], hdr["NAXIS3"], hdr["NAXIS4"]) == (M, L, 1, C)
    assert hdr["CTYPE1"] == "RA---SIN" and hdr["CTYPE2"] == "DEC--SIN"
    assert hdr["CDELT1"] == pytest.approx(-2 / 3600.0) and hdr["CRVAL2"] == -30.0
    assert hdr["CRVAL3"] == 1 and hdr["CTYPE4"] == "WAVELENGTH"
    assert data.dtype.byteorder == ">"
    np.testing.assert_array_equal(data[:, 0], img)
    # data starts on a block boundary right after the padded header
    assert raw[2880:2880 + img.itemsize] == img.reshape(-1)[:1].astype(img.dtype.newbyteorder(">")).tobytes()


def test_argument_checks_like_the_reference(tmp_path):
    name = os.path.join(tmp_path, "x.fits")
    with pytest.raises(Exception, match="Data size"):
        fits.save_to_fits_image(name, 4, 4, 1, 1, 2, 2, 0, 0, "I", 1, 1, 2, np.zeros(16, np.float32))
    with pytest.raises(Exception, match="float or double"):
        fits.save_to_fits_image(name, 2, 2, 1, 1, 1, 1, 0, 0, "I", 1, 1, 1, np.zeros(4, np.int32))


def test_polarization_codes_match_fits_table_28():
    assert [fits.POLARIZATION_CODES[k] for k in ["I", "Q", "U", "V", "RR", "LL", "RL", "LR",
                                                  "XX", "YY", "XY", "YX"]] == \
        [1, 2, 3, 4, -1, -2, -3, -4, -5, -6, -7, -8]
