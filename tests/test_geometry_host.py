"""The host geometry of blocks built from their grid (eilmer4_amd.geometry,
what e4hip_block_create_geometry receives) against the oracle's restatement
of compute_primary_cell_geometric_data (sfluidblock.d:650-842; properties.d):
bit-identical on the reference's ramp grids, a curvilinear 3D grid and a
curvilinear 2D grid.  (The oracle's geometry is itself pinned by closed forms
and the ramp case in test_oracle_geometry.py.)"""
import numpy as np
import pytest

import pyoracle as O
import ramp_case as R
from eilmer4_amd import geometry as G
from test_oracle_geometry import curvy_grid


def _compare(g, g_oracle=None):
    dims, ni, nj, nk, geom = G.block_geometry(g)
    b = O.Block.from_grid(O.Gas.ideal(), g if g_oracle is None else g_oracle)
    cells, faces = b.geometry()
    c2, f2 = G.split_geometry(dims, ni, nj, nk, geom)
    assert np.array_equal(cells, c2)
    assert len(faces) == len(f2)
    for a, b2 in zip(faces, f2):
        assert np.array_equal(a, b2)
    return geom


def test_ramp_grids_bitwise():
    for g in R.grids():
        _compare(g)


def test_curvilinear_3d_bitwise():
    _compare(curvy_grid(9, 7, 5, seed=5))


def test_curvilinear_2d_bitwise():
    g = curvy_grid(6, 5, 1, seed=6)[0].copy()
    g[..., 2] = 0.0
    _compare(g[..., :2], g)


def test_degenerate_cell_refused():
    g = curvy_grid(4, 3, 3, seed=7)[:, :, ::-1]   # vertices in left-handed order: negative volumes
    with pytest.raises(ValueError):
        G.block_geometry(g)
