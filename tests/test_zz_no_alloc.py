"""isa_minimize_shifts_noalloc (tables in the library's fixed device array) against isa_minimize_shifts
(tables in a caller workspace): identical outputs, on two calls in a row, the second smaller."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, empty, host, p, stream, workspace  # noqa: E402


def _problem(oracle, rng, n, strategy, block, ref, tiles):
    A, pairs, index, rows = oracle.shift_tables(n, ref, strategy, block)
    m = A.shape[0]
    x = np.concatenate([np.zeros((tiles, 1, 2)), rng.uniform(-6, 6, (tiles, n - 1, 2))], axis=1)
    d = x[:, rows[:, 1]] - x[:, rows[:, 0]] + rng.normal(0, 0.05, (tiles, m, 2))
    for t in range(0, tiles, 3):  # outliers: tiles iterate more than once
        d[t, rng.integers(0, m)] += rng.uniform(3, 8, 2)
    return np.ascontiguousarray(np.transpose(d, (1, 0, 2))).astype(np.float32)  # [m][tile][2]


def _run(isa, name, meas, n, strategy, block, ref, tiles, ws=None):
    import torch

    d_meas = dev(meas)
    r = empty((tiles, n - 1, 2), fill=-1.0)
    st = empty((tiles,), dtype=torch.int32, fill=-7)
    it = empty((tiles,), dtype=torch.int32, fill=-7)
    args = [p(d_meas), tiles * 8, p(r), p(st), p(it), tiles, n, ref, strategy, block]
    if ws is not None:
        args.append(p(ws))
    isa.call(name, *args, stream())
    return host(d_meas), host(r), host(st), host(it)


# the first call is the larger one: the second reads tables over the first call's leftovers
CALLS = [(24, 0, 5, 0, 53), (7, 3, 3, 2, 11)]


def test_noalloc_minimizer_matches_workspace_form(isa, oracle, gpu):
    rng = np.random.default_rng(2718)
    for n, strategy, block, ref, tiles in CALLS:
        print(f"noalloc minimizer: {n} frames, strategy {strategy}, {tiles} tiles", flush=True)
        meas = _problem(oracle, rng, n, strategy, block, ref, tiles)
        ws = workspace(isa.lib().isa_minimize_shifts_workspace(n, ref, strategy, block))
        want = _run(isa, "isa_minimize_shifts", meas, n, strategy, block, ref, tiles, ws)
        got = _run(isa, "isa_minimize_shifts_noalloc", meas, n, strategy, block, ref, tiles)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        assert np.all(got[2] == -1) and np.all(got[3] >= 1)


def test_noalloc_minimizer_rejects_oversized_tables(isa, gpu):
    meas = np.zeros((65 * 64 // 2, 1, 2), np.float32)
    with pytest.raises(isa.IsaError):
        _run(isa, "isa_minimize_shifts_noalloc", meas, 65, 0, 5, 0, 1)
