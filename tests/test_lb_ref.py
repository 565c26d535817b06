"""CPU checks of the learned-boundary reference (oracle/lb_ref.py): the
backward pass against central finite differences, bfloat16 rounding on
known values, the layer-cell partition and the feature normalisation."""
import numpy as np
import pytest

from oracle import lb_ref


def _rand_params(rng, L, scale=0.3):
    return rng.normal(0, scale, lb_ref.nparams(L))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    L, n = 3, 7
    p = _rand_params(rng, L)
    x = rng.normal(0, 1, (n, 81))
    d = rng.integers(0, L, n)
    t = rng.normal(0, 1, (n, 9))
    m = (rng.random((n, 9)) > 0.2).astype(float)
    _, _, g = lb_ref.loss_grad(p, L, x, d, t, m, rounded=False)
    idx = rng.choice(p.size, 60, replace=False)
    # every block of the layout is probed
    offs = np.cumsum([0, 81 * 64, L * 64, 64 * 64, 64, 64 * 9])
    idx = np.concatenate([idx, offs + 1])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (lb_ref.loss_grad(pp, L, x, d, t, m, rounded=False)[1] - lb_ref.loss_grad(pm, L, x, d, t, m, rounded=False)[1]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 + 1e-5 * abs(fd), (q, fd, g[q])


def test_zero_output_layer_gives_bias_only():
    rng = np.random.default_rng(4)
    L = 2
    w = lb_ref.unpack(_rand_params(rng, L), L)
    w["w3"][:] = 0
    w["b3"][:] = np.arange(9)
    y, _ = lb_ref.forward(lb_ref.pack(w), L, rng.normal(0, 1, (5, 81)), np.zeros(5, int))
    assert np.array_equal(y, np.tile(np.arange(9.0), (5, 1)))


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    e = np.float32(2.0 ** -8)   # bf16 spacing at 1 is 2^-7
    vals = np.array([one, one + e, one + 3 * e, one + np.float32(2 ** -7) + e, -2.5, 0.0], np.float32)
    got = lb_ref.bf16(vals)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 -> 1+2^-6 (nearest); 1+2^-7+2^-8 tie -> 1+2^-6 (even)
    assert got.tolist() == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -6, -2.5, 0.0]


@pytest.mark.parametrize("dims,L", [((12, 10, 9), 3), ((8, 8, 8), 4)])
def test_layer_cells_partition(dims, L):
    nx, ny, nz = dims
    cells = lb_ref.layer_cells(nx, ny, nz, L)
    inner = max(0, nx - 2 * L) * max(0, ny - 2 * L) * max(0, nz - 2 * L)
    assert len(cells) == nx * ny * nz - inner
    assert len({c[:3] for c in cells}) == len(cells)
    for i, j, k, f, d in cells:
        dd = (i, nx - 1 - i, j, ny - 1 - j, k, nz - 1 - k)
        assert d == min(dd) < L and dd[f] == d and all(dd[q] > d for q in range(f))


def test_features_are_rms_normalised_and_zero_outside():
    rng = np.random.default_rng(5)
    nx, ny, nz, L = 9, 8, 7, 2
    shapes = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz), (nx, ny, nz),
              (nx + 1, ny + 1, nz), (nx + 1, ny, nz + 1), (nx, ny + 1, nz + 1)]
    fields = [rng.normal(0, 1, s) for s in shapes]
    cells = lb_ref.layer_cells(nx, ny, nz, L)
    X, R, dep = lb_ref.features(fields, cells, nx, ny, nz, 0.5)
    assert np.allclose(np.sqrt((X ** 2).mean(1)), 1.0, atol=1e-12)
    # cell (0, 3, 3) on face x-: its -1 and -2 x-neighbours are outside
    s = [q for q, c in enumerate(cells) if c[:3] == (0, 3, 3)][0]
    assert cells[s][3] == 0 and np.all(X[s, 9:18] == 0) and np.all(X[s, 63:72] == 0)
    assert X[s, 0] * R[s] == fields[0][0, 3, 3] and np.isclose(X[s, 3] * R[s], 0.5 * fields[3][0, 3, 3])
