"""Raw ingest on the GPU (isa_raw_decode) against the oracle's restatement of the reference's decoders, bit for
bit, on the writer's files; batches of mixed files; the reference's behaviour on streams that end early, hold a
code that consumes no bits, or decode out of range; and full-size round trips."""
from __future__ import annotations

import struct

import numpy as np
import pytest

import raw_cases
import raw_writer as W

pytestmark = pytest.mark.gpu

CASES = raw_cases.cases()


@pytest.fixture(scope="module")
def R():
    from oracle import raw

    return raw


def gpu_decode(files):
    from imagestackalignator_amd import raw as P

    outs = P.decode_files(files)
    return [o.cpu().numpy().view(np.uint16) for o in outs]


@pytest.mark.parametrize("name", sorted(CASES))
def test_decode_matches_oracle(gpu, isa, R, name):
    from imagestackalignator_amd import raw as P

    data, expected = CASES[name]
    print(f"raw case {name}", flush=True)
    got = gpu_decode([P.RawFile(data)])[0]
    ref = R.RawFile(data).raw_image
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, expected)


def test_batch_of_mixed_files(gpu, isa, R):
    from imagestackalignator_amd import raw as P

    names = sorted(CASES)
    files = [P.RawFile(CASES[n][0]) for n in names]
    got = gpu_decode(files)
    for n, g in zip(names, got):
        np.testing.assert_array_equal(g, R.RawFile(CASES[n][0]).raw_image, err_msg=n)


def test_restart_with_predictor_2_as_reference(gpu, isa, R):
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(48, 64, 14, 9)
    data = W.write_dng(img, bits=16, white=65535, tile=(32, 16), comps=2, psv=2, restart_rows=4)
    np.testing.assert_array_equal(gpu_decode([P.RawFile(data)])[0], R.RawFile(data).raw_image)


def _pef_with_stream(img, stream: bytes, lengths):
    """A PEF whose strip is replaced by `stream` (same length or shorter, padded with zero bytes)."""
    data = bytearray(W.write_pef(img, lengths=lengths))
    enc, _ = W.pentax_encode(img, lengths)
    k = data.find(enc)
    assert k > 0
    data[k:k + len(enc)] = stream.ljust(len(enc), b"\x00")
    return bytes(data)


def test_code_that_consumes_nothing_repeats(gpu, isa, R):
    # A table that leaves the codes starting with 1111 unassigned: once the stream reaches such bits, the
    # reference's reader takes a zero-length category 0 again and again, so the rest of the image repeats the
    # last values of each column phase.
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(24, 40, 12, 2)
    lengths = {0: 2, 1: 2, 2: 3, 3: 4, 4: 5, 5: 6, 6: 7, 7: 8, 8: 9, 9: 10, 10: 11, 11: 12, 12: 12}
    enc, _ = W.pentax_encode(img, lengths)
    cut = len(enc) // 3
    data = _pef_with_stream(img, enc[:cut] + b"\xFF" * 8, lengths)
    ref = R.RawFile(data).raw_image
    got = gpu_decode([P.RawFile(data)])[0]
    np.testing.assert_array_equal(got, ref)
    assert (ref[-1] == np.tile(ref[-1, :2], 20)).all()  # the last row holds the two repeated phase values


def test_stream_past_end_of_file_is_an_error(gpu, isa, R):
    from imagestackalignator_amd import IsaError
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(24, 40, 12, 3)
    data = W.write_pef(img)
    # move the strip to the end of the file without slack and cut its last bytes off
    blob = bytearray(data)
    enc, _ = W.pentax_encode(img)
    k = blob.find(enc)
    cut = bytes(blob) + enc[: len(enc) - 6]
    o = "<"
    # point StripOffsets at the copy appended at the end
    ifd0 = struct.unpack_from(o + "I", cut, 4)[0]
    n = struct.unpack_from(o + "H", cut, ifd0)[0]
    b2 = bytearray(cut)
    for e in range(n):
        at = ifd0 + 2 + 12 * e
        if struct.unpack_from(o + "H", b2, at)[0] == 0x111:
            struct.pack_into(o + "I", b2, at + 8, len(blob))
    with pytest.raises(R.RawError):
        R.RawFile(bytes(b2))
    with pytest.raises(IsaError) as e:
        gpu_decode([P.RawFile(bytes(b2))])
    assert e.value.code == 5
    assert k > 0


def test_value_out_of_range_is_an_error(gpu, isa, R):
    from imagestackalignator_amd import IsaError
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(20, 24, 14)
    for psv in (1, 3):
        data = bytearray(W.write_dng(img, tile=(24, 20), comps=1, psv=psv))
        k = data.find(b"\xFF\xC3")
        data[k + 4] = 10  # declares 10-bit samples
        with pytest.raises(R.RawError):
            R.RawFile(bytes(data))
        with pytest.raises(IsaError) as e:
            gpu_decode([P.RawFile(bytes(data))])
        assert e.value.code == 5


@pytest.mark.parametrize("shape", [(768, 1024), (1531, 2047)])
def test_pef_round_trip_large(gpu, isa, shape):
    from imagestackalignator_amd import raw as P

    h, w = shape
    img = raw_cases.scene(h, w, 14, 11)
    print(f"pef round trip {shape}", flush=True)
    got = gpu_decode([P.RawFile(W.write_pef(img))])[0]
    np.testing.assert_array_equal(got, img)


def test_dng_tiles_round_trip_large(gpu, isa, R):
    from imagestackalignator_amd import raw as P

    img = raw_cases.scene(1000, 1536, 14, 12)
    data = W.write_dng(img, tile=(256, 256), comps=2)
    got = gpu_decode([P.RawFile(data)])[0]
    np.testing.assert_array_equal(got, img)
    data = W.write_dng(img, layout="strips", rows_per_strip=1000, comps=2)  # one strip of 1.5 M samples
    np.testing.assert_array_equal(gpu_decode([P.RawFile(data)])[0], img)


def test_full_size_pef_burst_round_trip(gpu, isa):
    # 2 frames of 6000 x 4000, decoded in one call
    from imagestackalignator_amd import raw as P

    imgs = [raw_cases.scene(4000, 6000, 14, 20 + i, noise=60) for i in range(2)]
    files = [P.RawFile(W.write_pef(im)) for im in imgs]
    got = gpu_decode(files)
    for g, im in zip(got, imgs):
        assert np.array_equal(g, im)
