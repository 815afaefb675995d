"""Full-face block connections in the oracle: the ghost-to-source mapping
against the reference's case tables (tests/golden/full_face_copy_map.json,
written by make_connection_map.py from full_face_copy.d:142-962), its
consistency, and decomposition invariance of the multi-block step."""
import json
import os

import numpy as np
import pytest

import pyoracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "full_face_copy_map.json")
FACES = ["north", "east", "south", "west", "top", "bottom"]
AXIS = {"i": 0, "j": 1, "k": 2}
NORMAL = {"north": 1, "south": 1, "east": 0, "west": 0, "top": 2, "bottom": 2}


def test_mapping_matches_reference_tables():
    d = json.load(open(GOLDEN))
    assert len(d["entries"]) == 16 + 144
    n = [5, 4, 3]
    checked = 0
    for e in d["entries"]:
        dims, f, g, o = e["dims"], e["face"], e["other_face"], e["orientation"]
        fax, gname = d["in_face_axes"][FACES[f]], FACES[g]
        nn = n if dims == 3 else [5, 4, 1]
        # destination positions that stay inside the source block under any mapping
        for layer in (1, 2):
            for a in range(3):
                for b in (range(3) if dims == 3 else [0]):
                    dest = {fax[0]: a}
                    if dims == 3:
                        dest[fax[1]] = b
                    src = O.connection_source(dims, f, g, o, nn, layer, a, b)
                    for sa, (dv, rev) in e["src"].items():
                        want = nn[AXIS[sa]] - dest[dv] - 1 if rev else dest[dv]
                        assert src[AXIS[sa]] == want, (e, layer, a, b, src)
                    dn = NORMAL[gname]
                    plus = gname in ("north", "east", "top")
                    assert src[dn] == (nn[dn] - layer if plus else layer - 1)
                    checked += 1
    assert checked > 1000


@pytest.mark.parametrize("dims", [2, 3])
def test_mapping_is_inverse_of_the_other_side(dims):
    """connectBlocks gives both sides the same orientation: mapping a face's
    ghost position across and back must return it (a dropped reversal or a
    swapped rotation fails here)."""
    ax = {0: (0, 2), 1: (1, 2), 2: (0, 2), 3: (1, 2), 4: (0, 1), 5: (0, 1)}
    nf, no, m = (6, 4, 4) if dims == 3 else (4, 1, 5)
    n = [m, m, m if dims == 3 else 1]
    for f in range(nf):
        for g in range(nf):
            for o in range(no):
                for a in range(m):
                    for b in (range(m) if dims == 3 else [0]):
                        s = O.connection_source(dims, f, g, o, n, 1, a, b)
                        a2, b2 = s[ax[g][0]], (s[ax[g][1]] if dims == 3 else 0)
                        t = O.connection_source(dims, g, f, o, n, 1, a2, b2)
                        assert (t[ax[f][0]], t[ax[f][1]] if dims == 3 else 0) == (a, b), (f, g, o, a, b)


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    T = 300.0 * (1.0 + 0.3 * rng.random(shape))
    v = [100.0 * (rng.random(shape) - 0.5) for _ in range(3)]
    return p, T, v


def test_decomposition_invariance_3d():
    """One 12x10x8 block against the same cells in two blocks split in i and
    joined east-west (and periodic in j through self-connections): identical
    after three steps (the connection-test pattern)."""
    gas = O.Gas.ideal()
    ni, nj, nk, cut = 12, 10, 8, 5
    p, T, v = _field((nk, nj, ni), 7)
    whole = O.Block(gas, 3, ni, nj, nk, 0.1, 0.1, 0.1)
    whole.set_config(flux_calc=O.AUSMDV)
    whole.set_from_pT(p, T, *v)
    whole.set_bc(O.NORTH, O.BC_PERIODIC)
    whole.set_bc(O.SOUTH, O.BC_PERIODIC)
    whole.set_bc(O.EAST, O.BC_EXTRAPOLATE)
    A = O.Block(gas, 3, cut, nj, nk, 0.1, 0.1, 0.1)
    B = O.Block(gas, 3, ni - cut, nj, nk, 0.1, 0.1, 0.1)
    for blk, sl in ((A, slice(0, cut)), (B, slice(cut, ni))):
        blk.set_config(flux_calc=O.AUSMDV)
        blk.set_from_pT(p[:, :, sl], T[:, :, sl], *[x[:, :, sl] for x in v])
        O.connect_blocks(blk, O.NORTH, blk, O.SOUTH, 0)
    O.connect_blocks(A, O.EAST, B, O.WEST, 0)
    B.set_bc(O.EAST, O.BC_EXTRAPOLATE)
    dt = 2.0e-5
    for _ in range(3):
        assert whole.step(dt)[0] == 0
        assert O.multi_step([A, B], dt)[0] == 0
    Uw = whole.get_U().reshape(5, nk, nj, ni)
    Ua = A.get_U().reshape(5, nk, nj, cut)
    Ub = B.get_U().reshape(5, nk, nj, ni - cut)
    assert np.array_equal(Uw[..., :cut], Ua)
    assert np.array_equal(Uw[..., cut:], Ub)


def test_rotated_partner_2d_one_stage():
    """2D: a block split in j whose upper part is stored rotated by 180
    degrees (both index axes reversed) and joined north-to-north.  From rest,
    one Euler stage sees only pressure differences, so each half equals the
    single block's cells (momentum with both components negated in the rotated
    half); this exercises the reversed 2D mapping.  (Later stages would need
    the velocity reorientation the reference leaves off by default.)"""
    gas = O.Gas.ideal()
    ni, nj, cut = 10, 12, 6
    rng = np.random.default_rng(3)
    p = 1.0e5 * (1.0 + 0.3 * rng.random((1, nj, ni)))
    T = 300.0 * (1.0 + 0.2 * rng.random((1, nj, ni)))
    zero = np.zeros_like(p)
    rot = lambda x: x[:, cut:][:, ::-1, ::-1]
    whole = O.Block(gas, 2, ni, nj, 1, 0.1, 0.1)
    A = O.Block(gas, 2, ni, cut, 1, 0.1, 0.1)
    B = O.Block(gas, 2, ni, nj - cut, 1, 0.1, 0.1)
    for blk in (whole, A, B):
        blk.set_config(update_scheme=O.EULER)
    whole.set_from_pT(p, T, zero, zero, zero)
    A.set_from_pT(p[:, :cut], T[:, :cut], zero[:, :cut], zero[:, :cut], zero[:, :cut])
    B.set_from_pT(rot(p), rot(T), zero[:, cut:], zero[:, cut:], zero[:, cut:])
    O.connect_blocks(A, O.NORTH, B, O.NORTH, 0)
    dt = 2.0e-5
    assert whole.step(dt)[0] == 0
    assert O.multi_step([A, B], dt)[0] == 0
    Uw = whole.get_U().reshape(5, nj, ni)
    Ua = A.get_U().reshape(5, cut, ni)
    Ub = B.get_U().reshape(5, nj - cut, ni)
    scale = np.abs(Uw).reshape(5, -1).max(axis=1)
    for q, sign in ((0, 1), (1, -1), (2, -1), (4, 1)):
        assert np.max(np.abs(Uw[q, :cut] - Ua[q])) <= 1e-13 * scale[q]
        assert np.max(np.abs(sign * Uw[q, cut:][::-1, ::-1] - Ub[q])) <= 1e-13 * scale[q], q
