"""The separable general kernels' tap rule, on the CPU in float32 arithmetic.

proj_general.h: when a bin's footprint is at most one voxel wide (half width h <= 1/2), it meets at most
two cells of an axis, i0 = floor(c + 1/2 - h) and i0 + 1, where c is the crossing in voxel-index units, and
a cell's weight is the overlap of [c - h, c + h] with [i - 1/2, i + 1/2].  The per-bin kernels scan every
cell in [floor(c - h - 1/2), ceil(c + h + 1/2)] and keep the positive overlaps; both must find the same taps
with the same weights.  The column-driven matched back projector inverts the rule: the column i is a cell of
a bin exactly when its crossing lies in [i - 3/2 + h, i + 1/2 + h).
"""
import numpy as np
import pytest

f32 = np.float32


def overlap(c, h, i):
    """min(c + h, i + 1/2) - max(c - h, i - 1/2) in float32, as the kernels evaluate it."""
    return np.minimum(f32(c) + f32(h), f32(i) + f32(0.5)) - np.maximum(f32(c) - f32(h), f32(i) - f32(0.5))


@pytest.mark.parametrize("h", [0.5, 0.433, 0.25, 0.1, 0.5 * np.cos(np.pi / 6)])
def test_two_cells_carry_every_positive_overlap(h):
    rng = np.random.default_rng(11)
    c = np.concatenate([rng.uniform(-40.0, 200.0, 20000),
                        # crossings on and next to the cell edges, where rounding decides
                        np.repeat(np.arange(-3, 60), 6) + np.tile([-1e-6, -1e-7, 0.0, 1e-7, 1e-6, 0.5], 63) - h])
    c = c.astype(np.float32)
    h = f32(h)
    i0 = np.floor(c + (f32(0.5) - h)).astype(np.int64)
    for k in range(len(c)):
        lo, hi = int(np.floor(c[k] - h - f32(0.5))), int(np.ceil(c[k] + h + f32(0.5)))
        cells = [i for i in range(lo, hi + 1) if overlap(c[k], h, i) > 0]
        assert set(cells) <= {int(i0[k]), int(i0[k]) + 1}, (c[k], h, cells, i0[k])
    # weights of the two cells sum to the footprint width where the footprint lies inside the axis
    w = np.maximum(0, overlap(c, h, i0)) + np.maximum(0, overlap(c, h, i0 + 1))
    np.testing.assert_allclose(w, 2 * h, rtol=0, atol=4e-5)


@pytest.mark.parametrize("au", [1.0, 0.81, 1.37, -0.9])
def test_inverted_rule_finds_every_bin_of_a_column(au):
    """For a column i, the bins iu0 with i in {i0(iu0), i0(iu0) + 1} all lie in the inverted interval
    [i - 3/2 + h, i + 1/2 + h) / Au widened by 0.01 bin (the kernel's crossing is one fma; here the
    product and the sum round separately, which the margin covers as well)."""
    rng = np.random.default_rng(12)
    h = f32(0.45)
    au = f32(au)
    for _ in range(200):
        x0 = f32(rng.uniform(-30.0, 30.0))
        i = int(rng.integers(0, 40))
        iu0 = np.arange(-200, 200)
        xi = (au * iu0.astype(np.float32) + x0).astype(np.float32)  # the kernels form this with one fma
        i0 = np.floor(xi + (f32(0.5) - h)).astype(np.int64)
        member = (i0 == i) | (i0 + 1 == i)
        lo, hi = (f32(i) - f32(1.5) + h - x0) / au, (f32(i) + f32(0.5) + h - x0) / au
        lo, hi = min(lo, hi), max(lo, hi)
        b0, b1 = int(np.ceil(lo - f32(0.01))), int(np.floor(hi + f32(0.01)))
        found = set(iu0[member].tolist())
        assert found <= set(range(b0, b1 + 1)), (i, x0, sorted(found), b0, b1)
