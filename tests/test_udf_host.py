"""Host side of the user-defined boundary and source terms (no GPU): the
ghost states UserDefinedBC builds and the source array built from a
vectorised sourceTerms(t, cell)."""
from types import SimpleNamespace

import numpy as np
import pytest

import eilmer4_amd as e4
import mms_case as mms


def _stub_block(gas, ni, nj, h):
    return SimpleNamespace(gas=gas, dims=2, ni=ni, nj=nj, nk=1, dx=h, dy=h, dz=1.0, id=0, nq=5, ncells=ni * nj)


@pytest.mark.parametrize("face", [e4.Face.south, e4.Face.west])
def test_user_defined_ghost_states(face):
    """The ghost-cell states of UserDefinedBC: p, T and velocity from the
    user's function at each ghost-cell centre, rho, u and a from the gas
    (user_defined_effects.d:179-226), S = 0, mass fraction 1 -- the same
    values, bit for bit, as the case's own construction of them."""
    gas = e4.IdealGas(mMass=mms.MMASS, gamma=mms.GAMMA)
    layout, h = mms.block_layout(16, 2, 2)
    ib, jb, ni, nj, origin = layout[3]
    blk = _stub_block(gas, ni, nj, h)
    bc = e4.bc.UserDefinedBC(lambda t, x, y, z: dict(zip(("p", "T", "velx", "vely"), mms.analytic_state(x, y)[1:])),
                             origin=origin)
    got = bc.fs(blk, face)
    want = mms.ghost_states_for_face(ni, nj, h, int(face), origin)
    assert bc.kind == e4.BCType.inflow_profile
    assert got.shape == want.shape == (2, ni, 10)
    assert np.array_equal(got, want)


def test_source_terms_array():
    """sourceTerms(t, cell) over the cell centres, missing entries 0, in the
    conserved order; the cell table carries x, y, z, i, j, k."""
    gas = e4.IdealGas()
    ni, nj, h = 6, 4, 0.25
    blk = _stub_block(gas, ni, nj, h)
    seen = {}

    def src(t, cell):
        seen.update(cell)
        return {"mass": cell["x"], "momentum_y": 2.0 * cell["y"], "total_energy": 7.0}

    Q = e4.udf.source_terms_array(blk, src, t=0.5, origin=(1.0, 2.0, 0.0))
    j, i = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    x = 1.0 + (i.ravel() + 0.5) * h
    y = 2.0 + (j.ravel() + 0.5) * h
    assert Q.shape == (5, ni * nj)
    assert np.array_equal(Q[0], x)
    assert np.array_equal(Q[1], np.zeros(ni * nj))
    assert np.array_equal(Q[2], 2.0 * y)
    assert np.array_equal(Q[3], np.zeros(ni * nj))
    assert np.array_equal(Q[4], np.full(ni * nj, 7.0))
    assert np.array_equal(seen["i"], i.ravel()) and np.array_equal(seen["j"], j.ravel())


def test_source_terms_unknown_species():
    gas = e4.ThermallyPerfectGas.from_json()
    blk = SimpleNamespace(gas=gas, dims=2, ni=2, nj=2, nk=1, dx=1.0, dy=1.0, dz=1.0, id=0,
                          nq=5 + gas.n_species, ncells=4)
    with pytest.raises(ValueError, match="not a valid species name"):
        e4.udf.source_terms_array(blk, lambda t, c: {"species": {"Xe": 1.0}})
    Q = e4.udf.source_terms_array(blk, lambda t, c: {"species": {gas.species[1]: 3.0}})
    assert np.array_equal(Q[6], np.full(4, 3.0)) and not Q[5].any()
