"""Structured fluid block on the device (the SFluidBlock role,
eilmer/sfluidblock.d:48-126): uniform Cartesian cells, or any structured grid
given by its vertices (SFluidBlock.from_grid)."""
import ctypes as C

import numpy as np

from . import lib as _lib
from .fvcore import BCType, Face


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class SFluidBlock:
    """A block of ni x nj (x nk) cells of size dx x dy (x dz) whose conserved
    state lives in device memory; every face starts as a slip wall."""

    def __init__(self, gas, dims, ni, nj, nk=1, dx=1.0, dy=1.0, dz=1.0, label="", geom=None):
        self.gas = gas
        self.dims = dims
        self.ni, self.nj, self.nk = ni, nj, (nk if dims == 3 else 1)
        self.dx, self.dy, self.dz = dx, dy, (dz if dims == 3 else 1.0)
        self.label = label
        self.geom = geom
        h = C.c_int(-1)
        if geom is None:
            _lib.check(_lib.load().e4hip_block_create(gas.handle(), dims, ni, nj, self.nk, dx, dy, self.dz,
                                                      C.byref(h)))
        else:
            g = np.ascontiguousarray(geom, dtype=np.float64)
            _lib.check(_lib.load().e4hip_block_create_geometry(gas.handle(), dims, ni, nj, self.nk, _p(g), 0,
                                                               C.byref(h)))
        self.id = h.value
        n = C.c_long(0)
        nq = C.c_int(0)
        _lib.check(_lib.load().e4hip_block_ncells(self.id, C.byref(n), C.byref(nq)))
        self.ncells, self.nq = n.value, nq.value
        self.bc = {f: BCType.slip_wall for f in Face}

    @classmethod
    def from_grid(cls, gas, grid, label=""):
        """A block of the structured grid with vertices grid [nk+1][nj+1][ni+1][3]
        (3D) or [nj+1][ni+1][2|3] (2D): cell volumes and lengths, face areas and
        frames from the vertices (eilmer4_amd.geometry)."""
        from .geometry import block_geometry
        dims, ni, nj, nk, geom = block_geometry(grid)
        return cls(gas, dims, ni, nj, nk, label=label, geom=geom)

    @property
    def shape(self):
        return (self.nk, self.nj, self.ni)

    def set_bc(self, face, kind, fs=None):
        """The boundary of one face: a BCType with its data, or one of the
        reference-named conditions of eilmer4_amd.bc (WallBC_WithSlip1(),
        OutFlowBC_Simple(), ...)."""
        from .bc import BoundaryCondition
        if isinstance(kind, BoundaryCondition):
            fs = kind.fs(self, face) if callable(kind.fs) else kind.fs
            kind = kind.kind
        fsv = None if fs is None else np.ascontiguousarray(fs, dtype=np.float64)
        nsp = getattr(self.gas, "n_species", 1)
        if kind == BCType.inflow_profile:
            # one state per ghost cell: [2][face cells][9 + nsp]
            d = 0 if Face(face) in (Face.east, Face.west) else (1 if Face(face) in (Face.north, Face.south) else 2)
            nfc = (self.nj if d == 0 else self.ni) * (self.nj if d == 2 else self.nk)
            if fsv is None or fsv.size != 2 * nfc * (9 + nsp):
                raise ValueError("a static profile needs %d ghost states of %d values" % (2 * nfc, 9 + nsp))
        elif kind == BCType.inflow and (fsv is None or fsv.size < 8 + (nsp if nsp > 1 else 0)):
            raise ValueError("an inflow needs a flow state")
        _lib.check(_lib.load().e4hip_block_set_bc(self.id, int(face), int(kind),
                                                  _p(fsv) if fsv is not None else None))
        self.bc[Face(face)] = BCType(kind)

    def set_axisymmetric(self, x0=0.0, y0=0.0):
        """2D axisymmetric geometry (GlobalConfig.axisymmetric) for this block,
        lower-left vertex at (x0, y0): per-radian volumes and face areas and
        the pressure source of the radial momentum."""
        _lib.check(_lib.load().e4hip_block_set_axisymmetric(self.id, float(x0), float(y0)))
        self.axisymmetric = (float(x0), float(y0))

    def set_source_terms(self, Q):
        """User-defined source terms (config.udf_source_terms): Q [nq][cells]
        in the conserved order (mass, x/y/z momentum, total energy, species
        densities), held constant until the next call; None removes them."""
        if Q is None:
            _lib.check(_lib.load().e4hip_block_set_source_terms(self.id, None))
            return
        Qv = np.ascontiguousarray(Q, dtype=np.float64)
        if Qv.size != self.nq * self.ncells:
            raise ValueError("source terms need [%d][%d] values" % (self.nq, self.ncells))
        _lib.check(_lib.load().e4hip_block_set_source_terms(self.id, _p(Qv)))

    def connect(self, face, other, other_face, orientation=0):
        """One side of a full-face connection: the ghost cells of `face` copy
        the mapped interior cells of `other` next to `other_face` at every
        stage (ExchangeBC_FullFace{otherBlock, otherFace, orientation},
        prep.lua:681-684; mapping full_face_copy.d:142-962)."""
        _lib.check(_lib.load().e4hip_block_connect(self.id, int(face), other.id, int(other_face), int(orientation)))
        self.bc[Face(face)] = BCType.connected

    def upload(self, U, Tguess=None):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nq, self.ncells)
        Tg = None if Tguess is None else np.ascontiguousarray(Tguess, dtype=np.float64).ravel()
        _lib.check(_lib.load().e4hip_block_upload(self.id, _p(U), _p(Tg) if Tg is not None else None))

    def set_from_pT(self, p, T, vx, vy, vz):
        """Initial state from pressure, temperature and velocity per cell
        (the initial-flow writer's FlowState(p, T), then encode_conserved)."""
        shp = self.shape
        p, T, vx, vy, vz = (np.broadcast_to(np.asarray(a, dtype=np.float64), shp).ravel() for a in (p, T, vx, vy, vz))
        rho, u, _ = self.gas.rho_u_a_from_pT(p, T)
        U = self.gas.encode(rho, u, vx, vy, vz)
        self.upload(U)
        return U

    def download(self, with_fs=False):
        U = np.empty((self.nq, self.ncells))
        nsp = getattr(self.gas, "n_species", 1)
        fs = np.empty((8 + (nsp if nsp > 1 else 0), self.ncells)) if with_fs else None
        _lib.check(_lib.load().e4hip_block_download(self.id, _p(U), _p(fs) if fs is not None else None))
        return (U, fs) if with_fs else U

    def destroy(self):
        if self.id >= 0:
            _lib.check(_lib.load().e4hip_block_destroy(self.id))
            self.id = -1


def connect_blocks(blkA, faceA, blkB, faceB, orientation=0):
    """Full-face connection of two blocks, both sides with the same
    orientation (connectBlocks, prep.lua:667-684)."""
    blkA.connect(faceA, blkB, faceB, orientation)
    blkB.connect(faceB, blkA, faceA, orientation)
