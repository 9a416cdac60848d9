"""Dense and ragged cells through the device binning and P2G, against the oracle.

The counting sort fills a cell's segment in arrival order; the segment sort then makes it ascending
(the reference's stable order, ParticleSolver.cpp:200-219 via the cell-index sort). Cells of up to 16
particles are sorted in registers, longer ones by insertion, so the cases here straddle that limit
(15, 16, 17) and go far past it (64, 300, 1000 particles in one cell), on a shuffled input so that
every segment arrives out of order. P2G then visits the dense cells in several batches of 8 per
colour round. Bars: binning bit-exact; P2G valid flags bit-exact and face velocities within a few
FP32 ulps of the velocity scale (same weights, another summation order).
"""
import numpy as np
import pytest

from banana_amd import scenes

pytestmark = pytest.mark.gpu

COUNTS = (1, 2, 3, 8, 9, 15, 16, 17, 31, 64, 300, 1000)


def dense_particles(n=24, seed=5):
    params, spec = scenes.dam_break(n)
    dims, bmin, dx = scenes.grid_dims(params)
    rng = np.random.default_rng(seed)
    # distinct interior cells (i, j, k), away from the dam's fluid block so their counts are exact
    cells = set()
    while len(cells) < len(COUNTS):
        cells.add(tuple(int(v) for v in rng.integers((dims[0] // 2 + 2, 3, 3), (dims[0] - 3, dims[1] - 3, dims[2] - 3))))
    cells = sorted(cells)
    pos = []
    for (i, j, k), c in zip(cells, COUNTS):
        u = rng.uniform(0.02, 0.98, (c, 3))
        pos.append((bmin.astype(np.float64) + (np.array([i, j, k]) + u) * np.float64(dx)).astype(np.float32))
    block = rng.uniform(0.05, 0.35, (4000, 3)).astype(np.float32)  # a sparse background in the low corner
    pos = np.concatenate(pos + [block])
    pos = pos[rng.permutation(pos.shape[0])]
    vel = rng.normal(0.0, 0.5, pos.shape).astype(np.float32)
    return params, dims, bmin, dx, cells, pos, vel


def test_dense_cells_binning_bit_exact(oracle, bnn):
    params, dims, bmin, dx, cells, pos, vel = dense_particles()
    g = bnn.FLIPSolver(params)
    g.set_particles(pos, vel)
    g.bin_particles()
    perm, start = g.bins()
    key, perm_o, start_o = oracle.bin_particles(pos, g.bmin, g.dx, g.n)
    assert np.array_equal(start, start_o)
    assert np.array_equal(perm, perm_o)
    counts = np.diff(start.astype(np.int64)).reshape(dims[2], dims[1], dims[0])
    got = sorted(int(counts[k, j, i]) for (i, j, k) in cells)
    assert got == sorted(COUNTS), got
    # within every cell the particles are in ascending input order
    for c in np.flatnonzero(np.diff(start.astype(np.int64)) > 1):
        seg = perm[start[c]:start[c + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0), c
    g.close()


def test_dense_cells_p2g(oracle, bnn):
    params, dims, bmin, dx, cells, pos, vel = dense_particles(seed=9)
    g = bnn.FLIPSolver(params)
    g.set_particles(pos, vel)
    o = oracle.Oracle(params)
    o.set_particles(pos, vel)
    g.map_particles_to_grid()
    o.map_particles_to_grid()
    for c, f, fv in (("u", oracle.F_U, oracle.F_VALID_U), ("v", oracle.F_V, oracle.F_VALID_V), ("w", oracle.F_W, oracle.F_VALID_W)):
        a = g.grid(c)
        b = o.field(f)
        assert np.array_equal(g.grid("valid_" + c), o.field(fv)), c
        assert np.max(np.abs(a - b)) <= 1e-5 * max(1.0, np.abs(b).max()), (c, float(np.max(np.abs(a - b))))
    g.close()
