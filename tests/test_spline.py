"""Equidistant-model interpolation (pmmc_spline_positions) against the
oracle's restatement of HierarchicalChromosome::gpuHelper's CPU path
(oracle/spline.py).  The reference's own path is in a .cu file and needs
CUDA, so the oracle is pinned by the literal loop, closed forms and hand
cases below; the GPU results must match it bit for bit."""
import numpy as np
import pytest

from oracle import spline as S


def chain(rng, n, gap=(200, 5000), start=10_000):
    lens = rng.integers(100, 2000, size=n)
    gaps = rng.integers(gap[0], gap[1], size=n)
    starts = start + np.cumsum(lens + gaps) - lens
    ends = starts + lens
    gp = ((starts + ends) // 2).astype(np.int64)
    xyz = np.cumsum(rng.normal(0, 1, size=(n, 3)), axis=0).astype(np.float32)
    return gp, xyz, int(ends[0]), int(starts[-1])


# ---------------- CPU: the oracle ----------------
@pytest.mark.parametrize("seed,n,res", [(0, 2, 50), (1, 3, 97), (2, 12, 333), (3, 40, 1000)])
def test_vectorised_oracle_equals_the_literal_loop(seed, n, res):
    rng = np.random.default_rng(seed)
    gp, xyz, fe, ls = chain(rng, n)
    start = int(gp[0]) - 3 * res              # before the first bead too
    n_out = S.count(start, int(gp[-1]) + 2 * res, res)
    a = S.scan_positions(gp, xyz, fe, ls, start, res, n_out)
    b = S.spline_positions(gp, xyz, fe, ls, start, res, n_out)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_interpolation_passes_through_beads_and_keeps_lines():
    gp = np.array([1000, 2000, 3000, 4000, 5000], np.int64)
    xyz = np.array([[0, 0, 0], [1, 2, 3], [2, 4, 6], [3, 6, 9], [4, 8, 12]], np.float32)
    # t = 0 at every interior bead: the blend returns the bead exactly
    out = S.spline_positions(gp, xyz, 1001, 4999, 2000, 1000, 3)
    assert np.array_equal(out, xyz[1:4])
    # a straight, evenly spaced chain: Catmull-Rom reproduces the line
    out = S.spline_positions(gp, xyz, 1001, 4999, 1500, 250, 13)
    want = ((1500 + 250 * np.arange(13) - 1000) / 1000.0)[:, None] * np.array([1, 2, 3])
    assert np.allclose(out, want, atol=1e-5)


def test_end_rules():
    gp = np.array([100, 200, 300], np.int64)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], np.float32)
    out = S.spline_positions(gp, xyz, first_end=150, last_start=280, pos_start=0, resolution=10,
                             n_out=40)
    assert np.array_equal(out[:15], np.repeat(xyz[:1], 15, 0))      # g < first_end
    assert np.array_equal(out[28:], np.repeat(xyz[2:], 12, 0))      # g >= last_start
    # between beads 0 and 1 the missing p0 is mirrored: 2*p1 - p2 with p1 = bead 0
    t = np.float32(50 / 100)
    b1 = np.float32(0.5) * ((-t ** 3 + 2 * t * t) - t)
    assert out[15][0] == pytest.approx(float(-1.0 * b1 + 0 + 1.0 * (0.5 * (-3 * t ** 3 + 4 * t * t + t)) + 1.0 * (0.5 * (t ** 3 - t * t))), abs=1e-6)
    assert S.count(0, 9, 5) == 3 and S.count(0, 10, 5) == 3 and S.count(5, 4, 1) == 0


# ---------------- GPU: the kernel ----------------
@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,res,over", [(0, 2, 50, 0), (5, 64, 113, 4), (6, 5000, 1000, 1),
                                             (7, 3, 1, 0)])
def test_spline_kernel_bit_exact(ctx, seed, n, res, over):
    rng = np.random.default_rng(seed)
    gp, xyz, fe, ls = chain(rng, n)
    start, stop = int(gp[0]), int(gp[-1])
    n_out = S.count(start, stop, res) + over
    print(f"spline n={n} res={res} n_out={n_out}", flush=True)
    got = ctx.spline_positions(gp, xyz, fe, ls, start, res, n_out=n_out)
    want = S.spline_positions(gp, xyz, fe, ls, start, res, n_out)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_spline_kernel_writes_every_output_beyond_one_grid(ctx):
    """17 M outputs: more than one sweep of the 65,536 x 256 grid."""
    rng = np.random.default_rng(11)
    gp, xyz, fe, ls = chain(rng, 2000, gap=(20_000, 40_000))
    start = int(gp[0]) - 5
    n_out = 17_000_000
    res = max(1, (int(gp[-1]) - start) // n_out)
    got = ctx.spline_positions(gp, xyz, fe, ls, start, res, n_out=n_out)
    want = S.spline_positions(gp, xyz, fe, ls, start, res, n_out)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_spline_rejects_unsorted_beads(ctx):
    from pmmc_amd.api import PmmcError
    with pytest.raises(PmmcError, match="sorted"):
        ctx.spline_positions([10, 5], np.zeros((2, 3), np.float32), 12, 4, 0, 1, n_out=3)
