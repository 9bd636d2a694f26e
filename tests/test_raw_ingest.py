"""Raw ingest on the CPU: the oracle's decoders against the independent writer (round trips), its metadata
restatement against values fixed by hand from the reference's rules, and the library's host-side container
parse (isa_raw_open / isa_raw_get_info, no GPU needed) against the oracle."""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np
import pytest

import raw_cases
import raw_writer as W

CASES = raw_cases.cases()


@pytest.fixture(scope="module")
def R():
    from oracle import raw

    return raw


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_decodes_writer_files(R, name):
    data, expected = CASES[name]
    f = R.RawFile(data)
    assert f.raw_image.shape == expected.shape
    np.testing.assert_array_equal(f.raw_image, expected)


def test_oracle_pef_levels_from_maker_note(R):
    img = raw_cases.scene(20, 24, 14)
    f = R.RawFile(W.write_pef(img, black=(510, 512, 514, 516), white_level=16000, white_point=(9000, 8192, 8192, 12000),
                              scaling=8192, cfa=(1, 0, 2, 1), roll=-3.5, orientation=6))
    # 4 black values in CFA order G R B G: red = 512, blue = 514, green = max(first G 510, last G 516)
    np.testing.assert_array_equal(f.black_level, np.float32([512, 516, 514]))
    np.testing.assert_array_equal(f.white_level, np.float32([16000 - 512, 16000 - 516, 16000 - 514]))
    np.testing.assert_array_equal(f.white_balance, np.float32([9000 / 8192, 1.0, 12000 / 8192]))
    assert f.bayer == [1, 0, 2, 1] and f.roll_angle == -3.5 and f.roll_angle_present and f.orientation == 6
    assert f.crop == (2, 2, 20, 16)
    g = R.RawFile(W.write_pef(img, black=None, white_level=None, bits=14))
    np.testing.assert_array_equal(g.white_level, np.float32([16384] * 3))  # 2^bits without a WhiteLevel tag
    np.testing.assert_array_equal(g.black_level, np.float32([0, 0, 0]))
    h = R.RawFile(W.write_pef(img, black=(100, 200, 300)))
    np.testing.assert_array_equal(h.black_level, np.float32([100, 200, 300]))


def test_oracle_dng_levels_crop_noise(R):
    img = raw_cases.scene(40, 60, 14)
    f = R.RawFile(W.write_dng(img, tile=(32, 16), black=(100, 120, 130, 90), cfa=(2, 1, 1, 0), white=15000,
                              neutral=((1, 2), (1, 1), (3, 4)), noise=(1e-4, 2e-6, 3e-4, 4e-6, 5e-4, 6e-6)))
    np.testing.assert_array_equal(f.black_level, np.float32([90, 130, 100]))  # R at site 3, B at site 0
    np.testing.assert_array_equal(f.white_level, np.float32([15000 - 90, 15000 - 130, 15000 - 100]))
    np.testing.assert_allclose(f.white_balance, [2.0, 1.0, 4 / 3], rtol=1e-7)
    assert (float(f.noise_alpha), float(f.noise_beta)) == (np.float32(3e-4), np.float32(4e-6))  # green plane
    g = R.RawFile(W.write_dng(img, tile=(32, 16), noise=(7e-4, 8e-6), active_area=(1, 3, 39, 57),
                              crop=(4, 9, 44, 30), plane_color=(0, 1, 2)))
    assert (float(g.noise_alpha), float(g.noise_beta)) == (np.float32(7e-4), np.float32(8e-6))
    # active area (top 1, left 3): both odd, so the CFA shifts down and right; the default crop then adds
    # origin[0] to both coordinates and uses size[0] for both sizes (the reference reads element 0 twice)
    assert g.bayer == [2, 1, 1, 0]  # RGGB shifted by one row and one column
    cl, ct, cw, ch = 3 + 4, 1 + 4, 44, 44
    assert g.crop == (cl, ct, cw - max(0, cl + cw - 58), ch - max(0, ct + ch - 38))
    h = R.RawFile(W.write_dng(img, tile=(32, 16), black=(10, 20, 30), plane_color=(2, 1, 0)))
    np.testing.assert_array_equal(h.black_level, np.float32([30, 20, 10]))  # planes B, G, R
    assert not h.color_twist_is_identity and h.color_twist[2, 0] == 1 and h.color_twist[0, 2] == 1


def test_oracle_dng_white_less_max_black_delta(R):
    img = raw_cases.scene(20, 40, 14)
    f = R.RawFile(W.write_dng(img, tile=(16, 16), black=(64,), white=16000, delta_h=np.arange(40) % 7,
                              delta_v=np.arange(20) % 5))
    assert f.max_black_delta == 6 + 4
    np.testing.assert_array_equal(f.white_level, np.float32([16000 - 64 - 10] * 3))


def test_oracle_version_picks_ssss16_rule(R):
    # an SSSS = 16 code carries no extra bits in DNG 1.2+ (-32768); a 1.1 file reads 16 more bits
    img = np.zeros((4, 8), np.uint16)
    img[:, 1::2] = 32768  # differences of 32768 modulo 2^16 along each row (and from the first prediction)
    f = R.RawFile(W.write_dng(img, bits=16, tile=(8, 4), comps=1, white=65535))
    np.testing.assert_array_equal(f.raw_image, img)
    g = R.RawFile(W.write_dng(img, bits=16, tile=(8, 4), comps=1, white=65535, version=(1, 1, 0, 0)))
    assert not np.array_equal(g.raw_image, img)  # 16 bits of the following codes were taken as the difference


def test_oracle_restart_with_predictor_2_follows_reference(R):
    # The reference applies predictors 2-7 from the second row of a tile on, also on the first row of a restart
    # interval, where T.81 uses the left neighbour: such a file differs from the writer's image exactly there.
    # (16-bit samples, so that the drifting values never exceed the stream's bit depth)
    img = raw_cases.scene(48, 64, 14, 9)
    f = R.RawFile(W.write_dng(img, bits=16, white=65535, tile=(32, 16), comps=2, psv=2, restart_rows=4))
    rows = f.raw_image != img
    assert rows.any() and not rows[0:4].any()  # rows of the first interval agree


def test_oracle_errors_where_reference_throws(R):
    img = raw_cases.scene(20, 24, 14)
    pef = W.write_pef(img)
    with pytest.raises(R.RawError):
        R.RawFile(b"XX" + pef[2:])
    with pytest.raises(R.RawError):
        R.RawFile(pef[:2] + b"\x2b\x00" + pef[4:])
    # strip rows beyond the image: 20 rows in strips of 8 -> the third strip runs past the last row
    with pytest.raises(R.RawError):
        R.RawFile(W.write_pef(img, compressed=False, rows_per_strip=8))
    # a 10-bit stream holding 14-bit values
    data = bytearray(W.write_dng(img, tile=(24, 20), comps=1))
    k = data.find(b"\xFF\xC3")
    data[k + 4] = 10
    with pytest.raises(R.RawError):
        R.RawFile(bytes(data))


def test_noise_model_interpolation():
    from imagestackalignator_amd.raw import noise_model_value

    nm = ([100, 200, 400], [1.0, 2.0, 4.0], [0.1, 0.2, 0.4])
    assert noise_model_value(nm, 50) == (1.0, 0.1)
    assert noise_model_value(nm, 400) == (4.0, 0.4)
    assert noise_model_value(nm, 800) == (4.0, 0.4)
    a, b = noise_model_value(nm, 300)  # weight (300 - 400) / (200 - 400) = 0.5
    assert a == pytest.approx(3.0) and b == pytest.approx(0.3)
    assert noise_model_value(([], [], []), 100) == (0.0, 0.0)


# ---- the library's container parse against the oracle (host code only) --------------------------------------

def _info_matches(P, R, data):
    p = P.RawFile(data)
    o = R.RawFile(data, decode=True)
    assert (p.raw_width, p.raw_height, p.bit_depth) == (o.width, o.height, o.bit_depth)
    assert p.bayer_pattern == list(o.bayer)
    np.testing.assert_array_equal(p.white_level, o.white_level)
    np.testing.assert_array_equal(p.black_level, o.black_level)
    np.testing.assert_array_equal(p.white_balance, o.white_balance)
    np.testing.assert_array_equal(p.color_twist, o.color_twist)
    assert p.color_twist_is_identity == o.color_twist_is_identity
    assert (p.crop_left, p.crop_top, p.cropped_width, p.cropped_height) == tuple(o.crop)
    assert (p.roll_angle, p.roll_angle_present) == (o.roll_angle, o.roll_angle_present)
    assert (p.noise_model_alpha, p.noise_model_beta) == (float(o.noise_alpha), float(o.noise_beta))
    assert p.iso == o.iso and p.exposure_time == tuple(o.exposure) and p.orientation == o.orientation
    assert p.make == o.make and p.unique_model_name == o.unique_model and p.recording_date == o.recording_date
    if o.is_dng:
        assert p.max_black_delta == o.max_black_delta


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_parse_matches_oracle(isa, R, name):
    from imagestackalignator_amd import raw as P

    _info_matches(P, R, CASES[name][0])


def test_library_parse_metadata_variants(isa, R):
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(40, 60, 14)
    for data in [
        W.write_pef(img, black=(510, 512, 514, 516), white_level=16000, cfa=(1, 0, 2, 1), roll=-3.5, orientation=6),
        W.write_pef(img, order=">", black=(100, 200, 300), roll=2.0, mn_header="PENTAX"),
        W.write_pef(img, black=None, white_level=None),
        W.write_dng(img, tile=(32, 16), black=(100, 120, 130, 90), cfa=(2, 1, 1, 0), white=15000,
                    noise=(1e-4, 2e-6, 3e-4, 4e-6, 5e-4, 6e-6)),
        W.write_dng(img, order=">", tile=(32, 16), noise=(7e-4, 8e-6), active_area=(1, 3, 39, 57), crop=(4, 9, 44, 30)),
        W.write_dng(img, tile=(32, 16), black=(10, 20, 30), plane_color=(2, 1, 0), black_type=W.RATIONAL),
        W.write_dng(img, tile=(32, 16), black=(64,), delta_h=np.arange(60) % 7, delta_v=np.arange(40) % 5),
        W.write_dng(img, tile=(32, 16), exif_iso=False, iso=800, exposure=(1, 4000), orientation=8),
    ]:
        _info_matches(P, R, data)


def test_library_open_errors(isa):
    from imagestackalignator_amd import IsaError
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(20, 24, 14)
    pef = W.write_pef(img)
    for bad in (b"XX" + pef[2:], pef[:2] + b"\x2b\x00" + pef[4:], pef[:6], W.write_pef(img, compressed=False,
                                                                                          rows_per_strip=8)):
        with pytest.raises(IsaError) as e:
            P.RawFile(bad)
        assert e.value.code == 5  # ISA_ERR_DATA
    with pytest.raises(IsaError):
        P.RawFile(pef, fmt="dng")  # no CFA IFD


def test_extra_camera_profile(isa, tmp_path):
    from imagestackalignator_amd import raw as P

    xml = tmp_path / "profiles.xml"
    xml.write_text("""<?xml version="1.0" encoding="utf-8"?>
<ExtraCameraProfiles xmlns:xsd="http://www.w3.org/2001/XMLSchema" xmlns:xsi="http://www.w3.org/2001/XMLSchema-instance">
  <Profiles>
    <ExtraCameraProfile>
      <Make>PENTAX</Make>
      <UniqueModel>PENTAX K-3</UniqueModel>
      <NoiseModel>
        <Iso><int>100</int><int>400</int></Iso>
        <Alpha><double>1e-4</double><double>4e-4</double></Alpha>
        <Beta><double>0.0</double><double>2e-6</double></Beta>
      </NoiseModel>
      <CropInfo><Top>3</Top><Left>1</Left><Width>50</Width><Height>30</Height></CropInfo>
    </ExtraCameraProfile>
  </Profiles>
</ExtraCameraProfiles>""")
    img = raw_cases.scene(36, 52, 14)
    f = P.RawFile(W.write_pef(img, iso=200))
    assert f.load_extra_camera_profile(xml)
    w = (200 - 400) / (100 - 400)
    assert f.noise_model_alpha == float(np.float32(w * 1e-4 + (1 - w) * 4e-4))
    assert f.noise_model_beta == float(np.float32((1 - w) * 2e-6))
    # crop: left max(2, 1) = 2, top 3, width 50 - max(0, 1 + 50 - (52 - 2)) = 49, height 30
    assert (f.crop_left, f.crop_top, f.cropped_width, f.cropped_height) == (2, 3, 49, 30)
    g = P.RawFile(W.write_pef(img, make="OTHER"))
    assert not g.load_extra_camera_profile(xml)
