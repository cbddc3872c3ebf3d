"""The bench's work counts (CPU only): the backprojection's voxels inside the
reconstruction circle and the deposits the Radon walks run, which its
roofline fractions divide by. Each is checked against an explicit loop over
the voxels, rows and z-block pairs."""
import numpy as np

import bench


def test_circle_voxels_counts_serial_mask():
    """SerialBackend.cpp:210-223: voxel (x, y) is reconstructed when
    (x - R)^2 + (y - R)^2 <= R^2, R = n/2 (integer division)."""
    for n in (1, 2, 7, 8, 33, 64):
        r = n // 2
        inside = sum(1 for y in range(n) for x in range(n)
                     if (x - r) * (x - r) + (y - r) * (y - r) <= r * r)
        assert bench.circle_voxels(n) == float(inside * n), n


def _walked_by_loops(vol, na, n_row):
    """Rows (columns) a pair of 8-slice z-blocks walks: those holding a pixel
    whose normalised value is nonzero (raw != 0 and raw != the slice
    minimum) in any of its 16 slices; a missing second z-block repeats the
    last one."""
    nz, h, w = vol.shape
    nzb = (nz + 7) // 8
    live = []
    for z in range(nz):
        s = vol[z]
        live.append((s != 0) & (s != s.min()))
    rows = cols = 0
    for zp in range((nzb + 1) // 2):
        blocks = [2 * zp, min(2 * zp + 1, nzb - 1)]
        mask = np.zeros((h, w), bool)
        for zb in blocks:
            for z in range(8 * zb, 8 * zb + 8):
                if z < nz:
                    mask |= live[z]
        rows += sum(1 for y in range(h) if mask[y].any())
        cols += sum(1 for x in range(w) if mask[:, x].any())
    return 16.0 * (rows * w * n_row + cols * h * (na - n_row))


def test_walked_deposits_sparse_volume():
    rng = np.random.default_rng(3)
    vol = np.zeros((24, 40, 36), np.float32)
    for _ in range(40):
        vol[rng.integers(24), rng.integers(40), rng.integers(36)] = rng.uniform(-1.0, 2.0)
    vol[5, 7, 9] = vol[5].min() if vol[5].min() < 0 else -3.0  # a pixel at the slice minimum
    got = bench.walked_deposits(vol, 90, 46)
    assert got["deposits"] == _walked_by_loops(vol, 90, 46)
    assert 0.0 < got["rows_walked_frac"] < 1.0 and 0.0 < got["columns_walked_frac"] < 1.0


def test_walked_deposits_dense_and_wide():
    rng = np.random.default_rng(4)
    dense = rng.uniform(0.5, 1.0, (32, 20, 20)).astype(np.float32)
    got = bench.walked_deposits(dense, 30, 16)
    # every row but the one holding only the slice minimum is walked; with
    # distinct random values each row has nonzero pixels
    assert got["deposits"] == 32.0 * 20 * 20 * 30
    # rows past 516 px take the 8-slice walk, which walks every row
    wide = np.zeros((8, 4, 600), np.float32)
    assert bench.walked_deposits(wide, 10, 5)["deposits"] == 8.0 * 4 * 600 * 10
