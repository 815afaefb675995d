"""Eilmer4's native structured-grid and flow-solution files, gziptext and
rawbinary, so that a run of the device path starts from the reference's own
input files and leaves files the reference's post-processing reads.

Grid files (geom/grid/sgrid.d:590-740): header "structured_grid 1.0", label,
dimensions, niv/njv/nkv, then the vertices x y z with i fastest.  Flow files
(sfluidblock.d:1778-1955): header "structured_grid_flow 1.0", label, sim_time,
the variable names, dimensions, nicell/njcell/nkcell, then one line (or one
record of doubles) per cell with i fastest, in the order of the variable list
(globalconfig.d:791-836; the values themselves fvcell.d:2374-2578).  A reader
takes the variables by name, as read_solution does when they are not in the
fixed order, so files with other variable sets (MHD, older names such as
"e[0]" and "T[0]") are read too.
"""
import gzip
import struct
from dataclasses import dataclass

import numpy as np

GRID_HEADER = "structured_grid 1.0"
FLOW_HEADER = "structured_grid_flow 1.0"


def _is_gzip(path):
    with open(path, "rb") as f:
        return f.read(2) == b"\x1f\x8b"


# ------------------------------------------------------------------ grids

@dataclass
class StructuredGrid:
    """Vertices [nkv][njv][niv][3] (i fastest in the files)."""
    vertices: np.ndarray
    dimensions: int
    label: str = ""

    @property
    def niv(self):
        return self.vertices.shape[2]

    @property
    def njv(self):
        return self.vertices.shape[1]

    @property
    def nkv(self):
        return self.vertices.shape[0]

    def cartesian_spacing(self, rtol=1.0e-9):
        """(dx, dy, dz, origin) when the grid is uniform and axis-aligned (the
        geometry the device path computes in closed form); ValueError
        otherwise."""
        v = self.vertices
        o = v[0, 0, 0]
        dx = (v[0, 0, -1, 0] - o[0]) / (self.niv - 1)
        dy = (v[0, -1, 0, 1] - o[1]) / (self.njv - 1)
        dz = (v[-1, 0, 0, 2] - o[2]) / (self.nkv - 1) if self.dimensions == 3 else 1.0
        want = cartesian_grid(self.dimensions, self.niv - 1, self.njv - 1, max(self.nkv - 1, 1), dx, dy, dz,
                              origin=tuple(o)).vertices
        scale = max(abs(dx), abs(dy), abs(dz) if self.dimensions == 3 else 0.0)
        if want.shape != v.shape or np.max(np.abs(v - want)) > rtol * scale:
            raise ValueError("the grid is not uniform and axis-aligned; the device path needs Cartesian cells")
        return dx, dy, dz, o.copy()


def cartesian_grid(dims, ni, nj, nk, dx, dy, dz=1.0, origin=(0.0, 0.0, 0.0), label=""):
    """The vertices of a uniform Cartesian block of ni x nj (x nk) cells."""
    nkv = nk + 1 if dims == 3 else 1
    x = origin[0] + dx * np.arange(ni + 1)
    y = origin[1] + dy * np.arange(nj + 1)
    z = origin[2] + dz * np.arange(nkv) if dims == 3 else np.full(1, origin[2])
    v = np.empty((nkv, nj + 1, ni + 1, 3))
    v[..., 0] = x[None, None, :]
    v[..., 1] = y[None, :, None]
    v[..., 2] = z[:, None, None]
    return StructuredGrid(v, dims, label)


def write_grid(path, grid, fmt="gziptext"):
    """StructuredGrid.write_to_gzip_file / write_to_raw_binary_file."""
    v = grid.vertices.reshape(-1, 3)
    if fmt == "rawbinary":
        lab = grid.label.encode()
        with open(path, "wb") as f:
            f.write(GRID_HEADER.encode())
            f.write(struct.pack("<i", len(lab)))
            f.write(lab)
            f.write(struct.pack("<4i", grid.dimensions, grid.niv, grid.njv, grid.nkv))
            f.write(np.ascontiguousarray(v, dtype="<f8").tobytes())
        return
    lines = [GRID_HEADER, "label: %s" % grid.label, "dimensions: %d" % grid.dimensions,
             "niv: %d" % grid.niv, "njv: %d" % grid.njv, "nkv: %d" % grid.nkv]
    lines += ["%.18e %.18e %.18e" % tuple(p) for p in v]
    with gzip.open(path, "wt") as f:
        f.write("\n".join(lines) + "\n")


def read_grid(path):
    """read_from_gzip_file / read_from_raw_binary_file (format from the file)."""
    if _is_gzip(path):
        with gzip.open(path, "rt") as f:
            lines = f.read().splitlines()
        if lines[0].split()[:1] != ["structured_grid"] or lines[0].split()[1] != "1.0":
            raise ValueError("grid format version found: %r" % lines[0])
        label = lines[1].partition(":")[2].strip()
        dims = int(lines[2].split(":")[1])
        niv, njv, nkv = (int(lines[3 + q].split(":")[1]) for q in range(3))
        data = np.array(" ".join(lines[6:6 + niv * njv * nkv]).split(), dtype=np.float64)
        return StructuredGrid(data.reshape(nkv, njv, niv, 3), dims, label)
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:len(GRID_HEADER)] != GRID_HEADER.encode():
        raise ValueError("unexpected grid header")
    off = len(GRID_HEADER)
    (nlab,) = struct.unpack_from("<i", buf, off)
    off += 4
    label = buf[off:off + nlab].decode()
    off += nlab
    dims, niv, njv, nkv = struct.unpack_from("<4i", buf, off)
    off += 16
    v = np.frombuffer(buf, dtype="<f8", count=niv * njv * nkv * 3, offset=off)
    return StructuredGrid(v.reshape(nkv, njv, niv, 3).copy(), dims, label)


# ------------------------------------------------------------------ flow

@dataclass
class FlowSolution:
    """One block's cell data: values[name] is an array over the cells, i
    fastest, for each variable of the file."""
    variables: list
    values: dict
    dimensions: int
    nic: int
    njc: int
    nkc: int
    sim_time: float = 0.0
    label: str = ""

    @property
    def ncells(self):
        return self.nic * self.njc * self.nkc

    def get(self, *names):
        """The first of the names the file holds (older files name u and T
        "e[0]" and "T[0]", and the conductivity "k[0]")."""
        for n in names:
            if n in self.values:
                return self.values[n]
        raise KeyError("none of %s in the flow file" % (names,))

    def massf(self, nsp):
        names = [v for v in self.variables if v.startswith("massf[")]
        order = sorted(names, key=lambda s: int(s[6:s.index("]")]))
        if len(order) != nsp:
            raise ValueError("the flow file holds %d mass fractions, the gas has %d" % (len(order), nsp))
        return np.stack([self.values[n] for n in order])


def flow_variable_list(gas):
    """GlobalConfig.build_flow_variable_list (globalconfig.d:791-836) for this
    path's configurations: no MHD, quality, radiation, energy modes or local
    time-stepping; species names with whitespace as dashes (massfName)."""
    names = ["pos.x", "pos.y", "pos.z", "volume", "rho", "vel.x", "vel.y", "vel.z", "p", "a", "mu", "k",
             "mu_t", "k_t", "S", "tke", "omega"]
    species = getattr(gas, "species", None) or [getattr(gas, "name", "air")]
    names += ["massf[%d]-%s" % (i, "-".join(s.split())) for i, s in enumerate(species)]
    if len(species) > 1:
        names.append("dt_chem")
    names += ["u", "T"]
    return names


def write_flow(path, flow, fmt="gziptext"):
    """SFluidBlock.write_solution (sfluidblock.d:1902-1955): S is written as an
    integer in text and as a double in rawbinary."""
    cols = [np.asarray(flow.values[n], dtype=np.float64).ravel() for n in flow.variables]
    if fmt == "rawbinary":
        lab = flow.label.encode()
        with open(path, "wb") as f:
            f.write(FLOW_HEADER.encode())
            f.write(struct.pack("<i", len(lab)))
            f.write(lab)
            f.write(struct.pack("<d", flow.sim_time))
            f.write(struct.pack("<i", len(flow.variables)))
            for n in flow.variables:
                f.write(struct.pack("<i", len(n)))
                f.write(n.encode())
            f.write(struct.pack("<4i", flow.dimensions, flow.nic, flow.njc, flow.nkc))
            f.write(np.ascontiguousarray(np.stack(cols, axis=1), dtype="<f8").tobytes())
        return
    head = [FLOW_HEADER, "label: %s" % flow.label, "sim_time: %.18e" % flow.sim_time,
            "variables: %d" % len(flow.variables), "".join(' "%s"' % n for n in flow.variables),
            "dimensions: %d" % flow.dimensions, "nicell: %d" % flow.nic, "njcell: %d" % flow.njc,
            "nkcell: %d" % flow.nkc]
    fmts = ["%d" if n == "S" else "%.18e" for n in flow.variables]
    rows = []
    for c in range(flow.ncells):
        rows.append(" " + " ".join(f % (int(col[c]) if f == "%d" else col[c]) for f, col in zip(fmts, cols)))
    with gzip.open(path, "wt") as f:
        f.write("\n".join(head + rows) + "\n")


def read_flow(path):
    """SFluidBlock.read_solution (sfluidblock.d:1778-1900), variables by name."""
    if _is_gzip(path):
        with gzip.open(path, "rt") as f:
            lines = f.read().splitlines()
        if lines[0].split() != FLOW_HEADER.split():
            raise ValueError("flow file format version found: %r" % lines[0])
        label = lines[1].partition(":")[2].strip()
        sim_time = float(lines[2].split(":")[1])
        nvar = int(lines[3].split(":")[1])
        names = [s.strip('"') for s in lines[4].split()]
        if len(names) != nvar:
            raise ValueError("variables: %d but %d names" % (nvar, len(names)))
        dims = int(lines[5].split(":")[1])
        nic, njc, nkc = (int(lines[6 + q].split(":")[1]) for q in range(3))
        n = nic * njc * nkc
        data = np.array(" ".join(lines[9:9 + n]).split(), dtype=np.float64).reshape(n, nvar)
    else:
        with open(path, "rb") as f:
            buf = f.read()
        if buf[:len(FLOW_HEADER)] != FLOW_HEADER.encode():
            raise ValueError("unexpected flow header")
        off = len(FLOW_HEADER)
        (nlab,) = struct.unpack_from("<i", buf, off); off += 4
        label = buf[off:off + nlab].decode(); off += nlab
        (sim_time,) = struct.unpack_from("<d", buf, off); off += 8
        (nvar,) = struct.unpack_from("<i", buf, off); off += 4
        names = []
        for _ in range(nvar):
            (m,) = struct.unpack_from("<i", buf, off); off += 4
            names.append(buf[off:off + m].decode()); off += m
        dims, nic, njc, nkc = struct.unpack_from("<4i", buf, off); off += 16
        n = nic * njc * nkc
        data = np.frombuffer(buf, dtype="<f8", count=n * nvar, offset=off).reshape(n, nvar).copy()
    values = {name: data[:, q].copy() for q, name in enumerate(names)}
    return FlowSolution(names, values, dims, nic, njc, nkc, sim_time, label)


# ------------------------------------------------- blocks <-> flow files

def cell_centres(dims, ni, nj, nk, dx, dy, dz=1.0, origin=(0.0, 0.0, 0.0)):
    """Cell-centre positions and volume of a uniform Cartesian block (what
    hex_cell_properties / xyplane_quad_cell_properties give for such cells;
    a 2D cell's volume is its area per unit depth)."""
    nk = nk if dims == 3 else 1
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    x = origin[0] + (i + 0.5) * dx
    y = origin[1] + (j + 0.5) * dy
    z = origin[2] + (k + 0.5) * dz if dims == 3 else np.full(i.shape, origin[2])
    vol = dx * dy * (dz if dims == 3 else 1.0)
    return x.ravel(), y.ravel(), z.ravel(), vol


def flow_from_block(block, sim_time=0.0, origin=(0.0, 0.0, 0.0), label="", S=None):
    """The flow solution of a device block (its downloaded flow state), in
    the variable list the reference writes for this gas."""
    gas = block.gas
    _, fs = block.download(with_fs=True)
    nsp = getattr(gas, "n_species", 1)
    x, y, z, vol = cell_centres(block.dims, block.ni, block.nj, block.nk, block.dx, block.dy, block.dz, origin)
    n = block.ncells
    zero = np.zeros(n)
    v = {"pos.x": x, "pos.y": y, "pos.z": z, "volume": np.full(n, vol), "rho": fs[0], "vel.x": fs[5],
         "vel.y": fs[6], "vel.z": fs[7], "p": fs[1], "a": fs[4], "mu": zero, "k": zero, "mu_t": zero,
         "k_t": zero, "S": zero if S is None else np.asarray(S, dtype=np.float64), "tke": zero,
         "omega": np.ones(n), "u": fs[3], "T": fs[2]}
    names = flow_variable_list(gas)
    mnames = [m for m in names if m.startswith("massf[")]
    if nsp > 1:
        for s, m in enumerate(mnames):
            v[m] = fs[8 + s]
        v["dt_chem"] = np.full(n, -1.0)
    else:
        v[mnames[0]] = np.ones(n)
    return FlowSolution(names, v, block.dims, block.ni, block.nj, block.nk if block.dims == 3 else 1,
                        sim_time, label)


def conserved_from_flow(flow, gas):
    """encode_conserved of the flow file's states (fvcell.d:494-560), as
    init_simulation does after read_solution; for a mixture also the
    temperature the first Newton solve starts from."""
    rho = flow.get("rho")
    u = flow.get("u", "e[0]")
    vel = [flow.get("vel.x"), flow.get("vel.y"), flow.get("vel.z")]
    nsp = getattr(gas, "n_species", 1)
    if nsp > 1:
        return gas.encode(rho, u, *vel, flow.massf(nsp)), flow.get("T", "T[0]").copy()
    return gas.encode(rho, u, *vel), None


# ------------------------------------------------------- static profiles

class FlowProfile:
    """The flow states of a profile file for InFlowBC_StaticProfile
    (flowstate.d:589-736): lines of cell data in the fixed flow-file order
    (scan_cell_data_from_fixed_order_string, fvcell.d:2580-), lines holding
    '#' skipped; each ghost cell takes the state of the profile point nearest
    to its centre under the match rule, the first such point on ties."""

    MATCHES = ("xyz-to-xyz", "xyA-to-xyA", "AyA-to-AyA", "xy-to-xR", "Ay-to-AR")

    def __init__(self, path, gas, match="xyz-to-xyz"):
        if match not in self.MATCHES:
            raise ValueError("Invalid match option: %r" % match)
        self.match = match
        self.names = flow_variable_list(gas)
        self.nsp = getattr(gas, "n_species", 1)
        rows = []
        opener = gzip.open if _is_gzip(path) else open
        with opener(path, "rt") as f:
            for line in f:
                if "#" in line or not line.strip():
                    continue
                rows.append([float(x) for x in line.split()[:len(self.names)]])
        self.data = np.array(rows, dtype=np.float64)
        col = {n: q for q, n in enumerate(self.names)}
        self.pos = self.data[:, [col["pos.x"], col["pos.y"], col["pos.z"]]]
        self.col = col

    def distance(self, my, other):
        """compute_distance (flowstate.d:641-689) from one position to many."""
        if self.match == "xyz-to-xyz":
            d = my[None, :] - other
            return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        if self.match == "xyA-to-xyA":
            dx, dy = my[0] - other[:, 0], my[1] - other[:, 1]
            return np.sqrt(dx ** 2 + dy ** 2)
        if self.match == "AyA-to-AyA":
            return np.abs(my[1] - other[:, 1])
        other_r = np.sqrt(other[:, 1] ** 2 + other[:, 2] ** 2)
        my_r = np.sqrt(my[1] ** 2 + my[2] ** 2)
        dr = my_r - other_r
        if self.match == "xy-to-xR":
            dx = my[0] - other[:, 0]
            return np.sqrt(dx * dx + dr * dr)
        return np.abs(dr)

    def nearest(self, my):
        """find_nearest_profile_point: the first point of least distance."""
        return int(np.argmin(self.distance(np.asarray(my, dtype=np.float64), self.pos)))

    def state(self, my):
        """[rho, p, T, u, a, vx, vy, vz, S, Y...] of the nearest point, with
        adjust_velocity (flowstate.d:717-735) for the radial matches."""
        row = self.data[self.nearest(my)]
        c = self.col
        vx, vy, vz = row[c["vel.x"]], row[c["vel.y"]], row[c["vel.z"]]
        if self.match in ("xy-to-xR", "Ay-to-AR"):
            r = np.sqrt(my[1] ** 2 + my[2] ** 2)
            vel_yz = np.sqrt(vy ** 2 + vz ** 2)
            sign = -1.0 if vy < 0.0 else 1.0
            vy, vz = sign * vel_yz * my[1] / r, sign * vel_yz * my[2] / r
        Y = [row[c[n]] for n in self.names if n.startswith("massf[")]
        return np.array([row[c["rho"]], row[c["p"]], row[c["T"]], row[c["u"]], row[c["a"]], vx, vy, vz,
                         row[c["S"]]] + (Y if self.nsp > 1 else [1.0]), dtype=np.float64)


def ghost_centres(dims, ni, nj, nk, dx, dy, dz, face, origin=(0.0, 0.0, 0.0)):
    """Centres of the two ghost layers of a face of a uniform Cartesian block,
    [2][face cells][3], layer-major with the in-face index a + b * na
    (north/south: a = i, b = k; east/west: a = j, b = k; top/bottom: a = i,
    b = j)."""
    nk = nk if dims == 3 else 1
    d = 0 if face in (1, 3) else (1 if face in (0, 2) else 2)
    plus = face in (0, 1, 4)
    n = (ni, nj, nk)
    na = nj if d == 0 else ni
    nb = nj if d == 2 else nk
    out = np.empty((2, na * nb, 3))
    h = (dx, dy, dz)
    for layer in (1, 2):
        for b in range(nb):
            for a in range(na):
                idx = [0, 0, 0]
                idx[d] = n[d] - 1 + layer if plus else -layer
                if d == 0:
                    idx[1], idx[2] = a, b
                elif d == 1:
                    idx[0], idx[2] = a, b
                else:
                    idx[0], idx[1] = a, b
                c = [origin[q] + (idx[q] + 0.5) * h[q] for q in range(3)]
                if dims == 2:
                    c[2] = origin[2]
                out[layer - 1, a + b * na] = c
    return out


def profile_ghost_states(profile, block, face, origin=(0.0, 0.0, 0.0)):
    """The flow states InFlowBC_StaticProfile gives the ghost cells of a face
    of the block (E4HIP_BC_INFLOW_PROFILE layout [2][face cells][9 + nsp])."""
    g = ghost_centres(block.dims, block.ni, block.nj, block.nk, block.dx, block.dy, block.dz, face, origin)
    return np.stack([np.stack([profile.state(p) for p in layer]) for layer in g])


def write_profile(path, flow, cells):
    """A profile file: a '#' line of the variable names, then the fixed-order
    lines of the chosen cells of a flow solution (as the reference's sampling
    of a solution writes them)."""
    cols = [np.asarray(flow.values[n], dtype=np.float64).ravel() for n in flow.variables]
    with open(path, "w") as f:
        f.write("# " + " ".join(flow.variables) + "\n")
        for c in cells:
            f.write(" ".join(("%d" % int(col[c])) if n == "S" else ("%.18e" % col[c])
                             for n, col in zip(flow.variables, cols)) + "\n")


# ------------------------------------------------------ job directory layout

def file_ext(fmt):
    """gz for gziptext, bin for rawbinary (flowsolution.d:88-89)."""
    return "bin" if fmt == "rawbinary" else "gz"


def make_path_name(mytype, tindx):
    """fileutil.d:15-22: grid/t0003, flow/t0003."""
    return "%s/t%04d" % (mytype, tindx)


def make_file_name(mytype, job, blk_id, tindx, fmt="gziptext"):
    """fileutil.d:24-33: flow/t0003/job.flow.b0001.t0003.gz."""
    return "%s/t%04d/%s.%s.b%04d.t%04d.%s" % (mytype, tindx, job, mytype, blk_id, tindx, file_ext(fmt))


def append_times(job_dir, job, tindx, sim_time, dt_global):
    """The times file connecting tindx to the simulation time
    (simcore.d:678-683; its header from prep.lua:2017-2022)."""
    import os
    path = os.path.join(job_dir, "config", job + ".times")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write("# tindx sim_time dt_global\n")
        f.write("%04d %.18e %.18e\n" % (tindx, sim_time, dt_global))


def read_times(job_dir, job):
    """[(tindx, sim_time, dt_global)] of a times file."""
    import os
    out = []
    with open(os.path.join(job_dir, "config", job + ".times")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            t, s, d = line.split()[:3]
            out.append((int(t), float(s), float(d)))
    return out


def write_solution_files(job_dir, job, blocks, tindx, sim_time, dt_global, fmt="gziptext", origins=None,
                         with_grid=False):
    """simcore.d:637-684 for these blocks: one flow file per block under
    flow/tNNNN (and grid files when asked), then the times-file line."""
    import os
    for n, blk in enumerate(blocks):
        origin = origins[n] if origins is not None else (0.0, 0.0, 0.0)
        fl = flow_from_block(blk, sim_time, origin)
        path = os.path.join(job_dir, make_file_name("flow", job, n, tindx, fmt))
        os.makedirs(os.path.dirname(path), exist_ok=True)
        write_flow(path, fl, fmt)
        if with_grid:
            g = cartesian_grid(blk.dims, blk.ni, blk.nj, blk.nk, blk.dx, blk.dy, blk.dz, origin)
            gpath = os.path.join(job_dir, make_file_name("grid", job, n, tindx, fmt))
            os.makedirs(os.path.dirname(gpath), exist_ok=True)
            write_grid(gpath, g, fmt)
    append_times(job_dir, job, tindx, sim_time, dt_global)
