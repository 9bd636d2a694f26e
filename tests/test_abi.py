"""CPU-side checks of the C ABI: the library loads, exports every function the
header declares, and its host-only entry points (parameter defaults, Gaussian
taps, shift-collection tables) agree with the oracle. No GPU needed."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "isa_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isa_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(isa):
    declared = _declared()
    assert len(declared) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", str(isa.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(isa_\w+)", out))
    missing = [d for d in declared if d not in exported]
    assert not missing, missing
    lib = isa.lib()
    for d in declared:
        assert hasattr(lib, d)


def test_default_params_match_reference_defaults(isa, oracle):
    p = isa.default_params()
    assert (p.sigma_debayer_tracking, p.sigma_debayer_accumulation) == (0.5, 0.5)
    assert abs(p.high_pass - 0.01) < 1e-9 and abs(p.high_pass_sigma - 0.0025) < 1e-9
    assert (p.n_levels, p.resize_level[0], p.tile_size[0], p.max_shift[0]) == (1, 1, 32, 2)
    assert (p.iterations_lk, p.window_size_lk, p.erode_size) == (3, 11, 5)
    assert abs(p.min_det_lk - 0.01) < 1e-9 and p.threshold_m == 1e5
    assert list(p.cfa) == [0, 1, 1, 2] and p.reference_index == -1
    # the oracle's defaults fill the same struct identically
    q = isa.IsaParams()
    oracle.lib().ora_default_params(C.addressof(q))
    assert bytes(q) == bytes(p)


@pytest.mark.parametrize("sigma", [0.0, 0.3, 0.5, 0.9, 1.0, 1.7, 3.0, 61.0])
def test_gaussian_taps(isa, oracle, sigma):
    buf = (C.c_float * 99)()
    n = isa.lib().isa_gaussian_taps(sigma, buf, 99)
    ref = oracle.gaussian_taps(sigma)
    assert n == len(ref)
    np.testing.assert_array_equal(np.array(buf[:n], np.float32), ref)
    if sigma > 0:
        # closed form: normalised exp(-x^2 / 2 sigma^2) with (int)(sigma/0.6 - 0.4)*2 + 3 taps
        size = min(int(sigma / 0.6 - 0.4) * 2 + 3, 99)
        x = np.arange(size) - size // 2
        g = np.exp(-x * x / (2.0 * sigma * sigma))
        np.testing.assert_allclose(ref, g / g.sum(), rtol=1e-6)


@pytest.mark.parametrize("n,ref,strategy,block", [
    (2, 0, 0, 5), (5, 0, 0, 5), (20, 0, 0, 5), (50, 0, 0, 5), (6, 2, 1, 5), (10, 0, 1, 5),
    (10, 3, 2, 4), (10, 8, 2, 4), (12, 6, 2, 5), (9, 0, 3, 3), (20, 0, 3, 5), (7, 0, 3, 9)])
def test_shift_tables_match_oracle(isa, oracle, n, ref, strategy, block):
    A, pairs, index, rows = oracle.shift_tables(n, ref, strategy, block)
    m = A.shape[0]
    pp = (C.c_int * (2 * m))()
    ii = (C.c_int * (n * n))()
    rr = (C.c_int * (2 * m))()
    assert isa.lib().isa_shift_tables(n, ref, strategy, block, pp, ii, rr) == m
    np.testing.assert_array_equal(np.array(pp[:]).reshape(m, 2), pairs)
    np.testing.assert_array_equal(np.array(ii[:]).reshape(n, n), index)
    np.testing.assert_array_equal(np.array(rr[:]).reshape(m, 2), rows)
    # every row of A is the interval indicator of its (a, b) endpoints with the right sign
    for k in range(m):
        a, b = rows[k]
        e = np.zeros(n - 1)
        lo, hi = min(a, b), max(a, b)
        e[lo:hi] = 1 if b > a else -1
        np.testing.assert_array_equal(A[k], e)


def test_shift_table_counts():
    from oracle import oracle as O
    # Full: n(n-1)/2; OnlyOnReference: n-1
    for n in (3, 8, 20, 30, 50):
        assert O.shift_tables(n, 0, 0, 5)[0].shape[0] == n * (n - 1) // 2
        assert O.shift_tables(n, 0, 1, 5)[0].shape[0] == n - 1


def test_errors_are_reported_without_a_gpu(isa):
    with pytest.raises(isa.IsaError) as e:
        isa.call("isa_resize_box", None, 0, None, 0, 0, 0, 1, None)
    assert e.value.code == 1 and "isa_resize_box" in str(e.value)
    assert isa.lib().isa_shift_tables(1, 0, 0, 5, None, None, None) < 0
    assert isa.lib().isa_shift_tables(5, 0, 7, 5, None, None, None) < 0


@pytest.mark.parametrize("h,w", [(1, 1), (5, 7), (64, 33)])
def test_tiff16_encode_matches_oracle(isa, oracle, tmp_path, h, w):
    """isa_tiff16_encode / isa_write_tiff16 (host code) write ImageFileDirectory.SaveAsTiff's bytes: the
    oracle's restatement, byte for byte, in memory and in a file."""
    from imagestackalignator_amd import export

    rng = np.random.default_rng(h * w)
    rgb = rng.integers(0, 65536, (h, w, 3)).astype(np.uint16)
    want = oracle.tiff16_bytes(rgb)
    assert export.tiff16_bytes(rgb) == want
    path = tmp_path / "out.tif"
    export.write_tiff16(str(path), rgb)
    assert path.read_bytes() == want
    assert isa.lib().isa_tiff16_encode(None, 0, rgb.ctypes.data, w, h) == len(want)
    small = (C.c_ubyte * 16)()
    assert isa.lib().isa_tiff16_encode(small, 16, rgb.ctypes.data, w, h) == 0  # capacity too small
    assert isa.lib().isa_tiff16_encode(None, 0, rgb.ctypes.data, 0, h) == 0
    with pytest.raises(isa.IsaError):
        isa.call("isa_write_tiff16", str(tmp_path / "no_such_dir" / "x.tif").encode(), rgb.ctypes.data, w, h)
