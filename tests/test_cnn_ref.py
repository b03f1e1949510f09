"""The CPU reference of the learned psi-extrapolation network (tests/cnn_ref.py), pinned by
central finite differences in float64 (its backward formulas) and by the bfloat16 rounding rule
(round to nearest, ties to even)."""
import numpy as np

import cnn_ref as R


def test_reference_backward_against_finite_differences():
    """pins the CPU reference's backward formulas (float64, no rounding)"""
    rng = np.random.default_rng(2)
    N, H, W = 1, 4, 5
    x = rng.standard_normal((N, H, W, 32))
    ws = [rng.standard_normal((32, 9, 32)) * 0.2 for _ in range(4)]
    bs = [rng.standard_normal(32) * 0.1 for _ in range(4)]
    tgt = rng.standard_normal((N, H, W, 2))
    loss, dws, dbs, _ = R.network(x, ws, bs, tgt, bf=False)
    h = 1e-6
    for (l, idx) in [(0, (3, 4, 1)), (1, (7, 0, 30)), (2, (0, 8, 5)), (3, (1, 4, 9)), (3, (0, 2, 2))]:
        wp = [w.copy() for w in ws]
        wm = [w.copy() for w in ws]
        wp[l][idx] += h
        wm[l][idx] -= h
        fd = (R.network(x, wp, bs, tgt, bf=False)[0] - R.network(x, wm, bs, tgt, bf=False)[0]) / (2 * h)
        assert abs(fd - dws[l][idx]) <= 1e-6 * max(1.0, abs(fd)), (l, idx, fd, dws[l][idx])
    for l in range(4):
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += h
        bm[l][1] -= h
        fd = (R.network(x, ws, bp, tgt, bf=False)[0] - R.network(x, ws, bm, tgt, bf=False)[0]) / (2 * h)
        assert abs(fd - dbs[l][1]) <= 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even gives 1; 1 + 3 * 2^-8 rounds up
    assert R.bf16(1.0 + 2.0 ** -8) == 1.0
    assert R.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert R.bf16(-2.5) == -2.5
    assert R.from_bits(R.bf16_bits(3.140625)) == 3.140625
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.all(np.abs(R.bf16(x) - x) <= 2.0 ** -8 * np.abs(x))
