"""User-defined source terms (config.udf_source_terms) for device blocks.

The reference calls a Lua function sourceTerms(t, cell) for every cell at
every stage and adds the returned table to the cell's source vector
(user_defined_source_terms.d:23-106, called from simcore.d:2445-2464): mass,
momentum_x/y/z, total_energy, a species table by name, and for a
single-species gas the mass term again as the species term.  Here the function
is evaluated on the host over the block's cell centres, vectorised, and the
values are held on the device (e4hip_block_set_source_terms) until they are set
again, so a source that depends on time is refreshed by the caller between
steps and one that depends on the cell's flow state is not available.
"""
import numpy as np

from .flowfiles import cell_centres


def source_terms_array(block, source_terms, t=0.0, origin=(0.0, 0.0, 0.0)):
    """Q [nq][cells] of a block from source_terms(t, cell), where cell is a
    dict of arrays over the block's cells (x, y, z, vol, i, j, k, blkId) and
    the result a dict of arrays or scalars (missing entries are 0)."""
    x, y, z, vol = cell_centres(block.dims, block.ni, block.nj, block.nk, block.dx, block.dy, block.dz, origin)
    nk = block.nk if block.dims == 3 else 1
    k, j, i = (a.ravel() for a in np.meshgrid(np.arange(nk), np.arange(block.nj), np.arange(block.ni),
                                              indexing="ij"))
    cell = {"x": x, "y": y, "z": z, "vol": vol, "i": i, "j": j, "k": k, "blkId": block.id}
    src = source_terms(t, cell)
    n = block.ncells
    Q = np.zeros((block.nq, n))

    def get(name):
        return np.broadcast_to(np.asarray(src.get(name, 0.0), dtype=np.float64), (n,))

    Q[0] = get("mass")
    Q[1] = get("momentum_x")
    Q[2] = get("momentum_y")
    Q[3] = get("momentum_z")
    Q[4] = get("total_energy")
    nsp = getattr(block.gas, "n_species", 1)
    if nsp > 1:
        species = src.get("species")
        if species is not None:
            names = list(block.gas.species)
            for name, val in species.items():
                if name not in names:
                    raise ValueError("In the user-defined source terms, the species name '%s' "
                                     "in the species table is not a valid species name." % name)
                Q[5 + names.index(name)] += np.broadcast_to(np.asarray(val, dtype=np.float64), (n,))
    return Q


def set_source_terms(block, source_terms, t=0.0, origin=(0.0, 0.0, 0.0)):
    """Evaluate source_terms over the block's cells and hold the values on the
    device; returns the array."""
    Q = source_terms_array(block, source_terms, t, origin)
    block.set_source_terms(Q)
    return Q
