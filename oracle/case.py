"""CPU restatement of PIRO's case driver and Paraview writer.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of
piro_case_run and piro_paraview_write; the product never imports it).

Restated from the reference, function by function:

  read_ini            file_utilities::IniReader        src/fileutilities.cpp:127-176
  split_string etc.   string_utilities::*               src/stringutilities.cpp:9-134
  setup               Piro::preprocess + bc::readbc     src/preprocess.cpp:20-205,
                      / initbc / setbc / opencl_initBC  src/bc.cpp:20-36, 67-78, 107-165
  initial_condition   Piro::initialcondition            src/ic.cpp:16-99
  run                 Piro::solve + process::* +        src/solve.cpp:11-48,
                      CLBuffer::operator= + operator& + src/process.cpp:52-249,
                      scalarMatrix::Solve               src/operatoroverload.cpp:20-100,309-329
  writevti/writevth   paraview::writevti/writevth       src/paraview.cpp:12-137

The arithmetic of every step goes through the C oracle (oracle/__init__.py).
Deliberate departures, shared with the product and recorded in its
documentation: the diffusivity row scaling is applied on the CSR path
(alpha_row_scaled, the reference computes it but never applies it), and
Solver 27 solves (I - dt L) x = b (here in FP64) where the reference
eliminates L once and returns a zero field.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import (CD, CSR, div_csr, face_indices, fe_csr_step, fe_dense_step, fe_ht_step, geam3, grad_csr,
               grad_norm, ht_build, ht_table_size, ic_gaussian, jli, lap_csr, lap_dense, lap_norm,
               lu_factor_f64, lu_solve_f64, row_scale, spmv_allrows_step)

_NUM = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|[+-]?(?:inf|nan))", re.IGNORECASE)
_INT = re.compile(r"\s*([+-]?\d+)")


# ------------------------------------------------------------------ text

def read_ini(path: str) -> dict:
    """section -> key -> value; trims spaces only, skips ';'/'#' lines, a
    missing file gives an empty table (src/fileutilities.cpp:127-163)"""
    data: dict = {}
    if not os.path.exists(path):
        return data
    section = ""
    with open(path) as f:
        for raw in f.read().split("\n"):
            line = raw.strip(" ")
            if not line or line[0] in ";#":
                continue
            if line[0] == "[" and line[-1] == "]":
                section = line[1:-1]
            else:
                p = line.find("=")
                if p >= 0:
                    data.setdefault(section, {})[line[:p].strip(" ")] = line[p + 1:].strip(" ")
    return data


def ini_get(data: dict, section: str, key: str, default: str = "default_value") -> str:
    return data.get(section, {}).get(key, default)


def split_string(s: str, delim: str = " ") -> list:
    """every delimiter splits, so repeated delimiters give empty tokens"""
    return s.split(delim)


def to_floats(tokens) -> list:
    """istringstream >> float per token, failures skipped; the value is
    rounded to float32 (strtof)"""
    out = []
    for t in tokens:
        m = _NUM.match(t)
        if m:
            out.append(np.float32(float(m.group(1))))
    return out


def to_ints(tokens) -> list:
    out = []
    for t in tokens:
        m = _INT.match(t)
        if m:
            out.append(int(m.group(1)))
    return out


def stoi(s: str) -> int:
    m = _INT.match(s)
    if not m:
        raise ValueError(f"stoi: {s!r}")
    return int(m.group(1))


def stof(s: str) -> np.float32:
    m = _NUM.match(s)
    if not m:
        raise ValueError(f"stof: {s!r}")
    return np.float32(float(m.group(1)))


def count_spaces(s: str) -> int:
    return s.count(" ")


def count_words(s: str) -> int:
    return len(s.split())


# ------------------------------------------------------------ case setup

def face_lists(n) -> list:
    """the six face index lists of bc::initbc (src/bc.cpp:107-146), edges
    and corners excluded: 0 x = nx-1, 1 x = 0, 2 y = ny-1, 3 y = 0,
    4 z = 0, 5 z = nz-1, each in increasing index order"""
    nx, ny, nz = n
    ind = np.arange(nx * ny * nz)
    kd, jd, i = ind // (ny * nx), (ind % (ny * nx)) // nx, ind % nx
    xi = (i != 0) & (i != nx - 1)
    yi = (jd != 0) & (jd != ny - 1)
    zi = (kd != 0) & (kd != nz - 1)
    return [ind[(i == nx - 1) & yi & zi], ind[(i == 0) & yi & zi], ind[(jd == ny - 1) & xi & zi],
            ind[(jd == 0) & xi & zi], ind[(kd == 0) & xi & yi], ind[(kd == nz - 1) & xi & yi]]


@dataclass
class Slot:
    name: str
    type: int                 # 0 constant, 1 scalar, 2 vector
    device: np.ndarray        # what the GPU buffer holds (N or 3N floats)
    host: np.ndarray          # the host CellData values (setbc may edit them)


@dataclass
class CaseSpec:
    casename: str
    simindex: int
    time_scheme: int
    space_scheme: int
    solver_scheme: int
    data_type: int
    n: tuple                  # padded dims
    l: tuple
    origin: tuple
    spacing: tuple
    dt: np.float32
    totaltime: np.float32
    save: int
    slots: list
    has_bc: bool
    snapshot: np.ndarray | None
    jli: dict = field(default_factory=dict)
    ht: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    def match(self, var: str) -> int:
        """process::matchscalartovar: first slot of that name, else 0"""
        for k, s in enumerate(self.slots):
            if s.name == var:
                return k
        return 0


def initial_condition(kind: str, n, l, ic: dict) -> np.ndarray:
    """src/ic.cpp:34-99 (an unknown kind returns no values)"""
    N = n[0] * n[1] * n[2]
    if kind == "Gaussian":
        scale = stof(ini_get(ic, "Gaussian", "Scalefactor"))
        sigma = to_floats(split_string(ini_get(ic, "Gaussian", "Sigma")))
        mean = to_floats(split_string(ini_get(ic, "Gaussian", "Median")))
        # values = scalefactor * calculategaussian(coord, mean, sigma)
        g = ic_gaussian(tuple(n), np.float32(l), np.float32(mean), np.float32(sigma), 1.0)
        return (np.float32(scale) * g).astype(np.float32)
    if kind == "Coulomb":
        center = [float(c) for c in to_floats(split_string(ini_get(ic, "Coulomb", "center")))]
        Z = float(stof(ini_get(ic, "Coulomb", "Z")))
        e = float(stof(ini_get(ic, "Coulomb", "e")))
        eps0 = float(stof(ini_get(ic, "Coulomb", "epsilon_0")))
        v = np.zeros(N, np.float32)
        for x in range(n[0]):
            for y in range(n[1]):
                for z in range(n[2]):
                    c = [np.float32(np.float32(i) * np.float32(ll) / np.float32(nn))
                         for i, ll, nn in ((x, l[0], n[0]), (y, l[1], n[1]), (z, l[2], n[2]))]
                    # calc_length_3d: float differences, std::pow in double (src/mathoperations.cpp:42-46)
                    r = np.float32(math.pow(sum(float(np.float32(c[d] - np.float32(center[d]))) ** 2
                                                for d in range(3)), 0.5))
                    v[x + y * n[0] + z * n[0] * n[1]] = np.float32(
                        -(Z * 1e15 * e ** 2) / (4 * math.pi * eps0 * (float(r) + 1e-10)))
        return v
    if kind == "UniformVector":
        d = to_floats(split_string(ini_get(ic, "Vector", "Direction")))
        val = to_floats(split_string(ini_get(ic, "Vector", "Value")))
        return np.concatenate([np.full(N, np.float32(d[k] * val[0]), np.float32) for k in range(3)])
    return np.zeros(0, np.float32)


def setup(assets_dir: str, case_name: str) -> CaseSpec:
    r = read_ini(os.path.join(assets_dir, case_name))
    nn = to_ints(split_string(ini_get(r, "Mesh", "n")))
    ll = to_floats(split_string(ini_get(r, "Mesh", "l")))
    oo = to_floats(split_string(ini_get(r, "Mesh", "o", "0 0 0")))
    ss = to_floats(split_string(ini_get(r, "Mesh", "s", "1 1 1")))
    n = tuple(v + 2 for v in nn[:3])
    N = n[0] * n[1] * n[2]
    ic = read_ini(os.path.join(assets_dir, "IC", "distribution.ini"))
    icfiles = split_string(ini_get(r, "IC", "filename"))
    cl = ini_get(r, "Simulation", "Constants")
    cnames, cvals = split_string(cl), to_floats(split_string(ini_get(r, "Simulation", "Values")))
    sl = ini_get(r, "Simulation", "Scalars")
    snames = split_string(sl)
    vl = ini_get(r, "Simulation", "Vectors", "")
    vnames = split_string(vl)
    slots = []
    for i in range(count_spaces(cl) + 1):
        v = np.full(N, cvals[i], np.float32)
        slots.append(Slot(cnames[i], 0, v.copy(), v.copy()))
    n_scal = count_spaces(sl) + 1
    for i in range(n_scal):
        v = initial_condition(icfiles[i], n, ll, ic)
        dev = np.zeros(N, np.float32)
        dev[: min(N, v.size)] = v[:N]
        slots.append(Slot(snames[i], 1, dev, v.copy()))
    for i in range(count_words(vl)):
        # every vector takes the IC entry after the scalars' (src/preprocess.cpp:160-171)
        v = initial_condition(icfiles[n_scal], n, ll, ic)
        dev = np.zeros(3 * N, np.float32)
        dev[: min(3 * N, v.size)] = v[: 3 * N]
        slots.append(Slot(vnames[i], 2, dev, v.copy()))
    for _ in range(3):  # the operator slots appended at the end of preprocess
        slots.append(Slot("", -1, np.zeros(0, np.float32), np.zeros(0, np.float32)))
    spec = CaseSpec(
        casename=ini_get(r, "File", "casename"), simindex=stoi(ini_get(r, "File", "simindex")),
        time_scheme=stoi(ini_get(r, "Schemes", "Time")), space_scheme=stoi(ini_get(r, "Schemes", "Space")),
        solver_scheme=stoi(ini_get(r, "Schemes", "Solver")), data_type=stoi(ini_get(r, "Solve", "Data")),
        n=n, l=tuple(ll[:3]), origin=tuple(int(v) for v in oo[:3]), spacing=tuple(int(v) for v in ss[:3]),
        dt=stof(ini_get(r, "Solve", "Timestep")), totaltime=stof(ini_get(r, "Solve", "TotalTime")),
        save=stoi(ini_get(r, "Solve", "Save")), slots=slots, has_bc=False, snapshot=None)
    if spec.solver_scheme == 40:
        rs = read_ini(os.path.join(assets_dir, "solver.ini"))
        spec.jli = {"tol": stof(ini_get(rs, "JLI", "tolerance")), "urf": stof(ini_get(rs, "JLI", "URF")),
                    "max_iter": stoi(ini_get(rs, "JLI", "max_iter"))}
    if spec.data_type == 2:
        rh = read_ini(os.path.join(assets_dir, "Data", "hashtable.ini"))
        spec.ht = {k: stof(ini_get(rh, "Table", k)) for k in ("a", "b", "c")}
        spec.ht["lf"] = stof(ini_get(rh, "Table", "LoadFactor"))
    # boundary conditions come from assets/setup.ini (bc::readbc)
    rb = read_ini(os.path.join(assets_dir, "setup.ini"))
    if count_spaces(ini_get(rb, "BC", "type")) > 1:
        props = split_string(ini_get(rb, "BC", "property"))
        vals = to_floats(split_string(ini_get(rb, "BC", "values")))
        faces = face_lists(n)
        for f in range(6):
            msv = spec.match(props[f])
            spec.slots[msv].host[faces[f]] = vals[f]
        spec.snapshot = spec.slots[1].host[:N].copy()
        spec.has_bc = True
    return spec


# ------------------------------------------------------------- time loop

def _operator(spec: CaseSpec, alpha_row_scaled: bool):
    n = spec.n
    if spec.simindex == 0:
        if spec.data_type == 1:
            # Piro::DENSE::laplacian: no diffusivity scaling (csrscale is CSR-only)
            return ("dense", lap_dense(*n, spec.space_scheme, lap_norm(spec.l[0], n[0] - 2)))
        A = lap_csr(*n, spec.space_scheme, lap_norm(spec.l[0], n[0] - 2))
        if spec.data_type == 0 and alpha_row_scaled:
            A = row_scale(A, spec.slots[spec.match("Alpha")].device[: spec.N])
        if spec.data_type == 2:
            ts = ht_table_size(A.nnz, spec.N, spec.ht["a"], spec.ht["b"], spec.ht["c"], spec.ht["lf"])
            keys, vals = ht_build(A, ts)
            return ("ht", keys, vals)
        return ("csr", A)
    if spec.simindex == 1:
        g = grad_norm(spec.l[0], n[0] - 2)
        G = [grad_csr(*n, spec.space_scheme, d, g) for d in range(3)]
        v3 = spec.slots[spec.match("v")].device
        return ("csr", geam3(G[0], G[1], G[2], v3))
    if spec.simindex == 2:
        # process::div("rho", "u") (src/process.cpp:168-212), CSR only
        if spec.data_type != 0:
            raise ValueError("div is CSR only")
        u3 = spec.slots[spec.match("u")].device
        return ("csr", div_csr(*n, spec.space_scheme, grad_norm(spec.l[0], n[0] - 2), u3))
    raise ValueError("simindex must be 0, 1 or 2")


def _ht_dense(spec, keys, vals):
    N = spec.N
    a = np.zeros((N, N))
    for k, v in zip(keys, vals):
        if k >= 0:
            a[k // N, k % N] = v
    return a


def run(spec: CaseSpec, max_steps: int = -1, alpha_row_scaled: bool = True):
    """returns (steps, field slot 1, exports) with exports a list of
    (ts, [(name, type, values printed)])"""
    N, dims = spec.N, spec.n
    op = _operator(spec, alpha_row_scaled)
    x = spec.slots[1].device[:N].copy()
    exports = []
    lu = None
    t = np.float32(0.0)
    steps = 0
    while t < spec.totaltime and (max_steps < 0 or steps < max_steps):
        if spec.time_scheme == 11:
            if op[0] == "ht":
                y = fe_ht_step(N, op[1], op[2], spec.dt, x)
            elif op[0] == "dense":
                y = fe_dense_step(op[1], spec.dt, x)
            elif spec.data_type == 3:
                y = spmv_allrows_step(op[1], spec.dt, x)
            else:
                y = fe_csr_step(dims, op[1], spec.dt, x, spec.snapshot if spec.has_bc else None)
        elif spec.time_scheme == 12 and spec.solver_scheme == 40:
            y, _, _ = jli(dims, op[1], spec.dt, x, spec.jli["urf"], spec.jli["tol"], spec.jli["max_iter"])
        elif spec.time_scheme == 12 and spec.solver_scheme == 27:
            if lu is None:
                L = _ht_dense(spec, op[1], op[2]) if op[0] == "ht" else op[1].dense().astype(np.float64)
                lu = lu_factor_f64(np.eye(N) - float(spec.dt) * L)
            y = lu_solve_f64(lu[0], lu[1], x.astype(np.float64)).astype(np.float32)
        else:
            raise ValueError("unsupported Time/Solver combination")
        # scalarMatrix::Solve: commit to slot 1, rewrite the faces, export
        x = y.astype(np.float32)
        if spec.has_bc:
            f = face_indices(*dims)
            x[f] = spec.snapshot[f]
        ts = int(np.float32(t / spec.dt))
        if (ts + 1) % spec.save == 0:
            printed = []
            for k, s in enumerate(spec.slots):
                if s.type in (1, 2):
                    printed.append((s.name, s.type, x if k == 1 else s.host))
            exports.append((ts, printed))
        t = np.float32(t + spec.dt)
        steps += 1
    return steps, x, exports


# -------------------------------------------------------------- Paraview

def fmt_float(v) -> str:
    """std::ostream << float: %g with 6 significant digits, glibc's
    inf/nan spelling (sign kept on NaN)"""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    return "%g" % v


def _interior(values, n, off=0):
    nx, ny, nz = n
    return " ".join(fmt_float(values[off + x + y * nx + z * nx * ny])
                    for z in range(1, nz - 1) for y in range(1, ny - 1) for x in range(1, nx - 1))


def writevti(n, origin, spacing, fields) -> str:
    """fields: (name, type, values) in slot order; type 1 scalar, 2 vector"""
    ext = f"0 {n[0] - 2} 0 {n[1] - 2} 0 {n[2] - 2}"
    s = '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">\n'
    s += (f'<ImageData WholeExtent="{ext}" Origin="{origin[0]} {origin[1]} {origin[2]}" '
          f'Spacing="{spacing[0]} {spacing[1]} {spacing[2]}">\n')
    s += f'<Piece Extent="{ext}">\n'
    s += '<PointData Scalars="ScalarValues">\n</PointData>\n<CellData>\n'
    N = n[0] * n[1] * n[2]
    has_vectors = any(t == 2 for _, t, _ in fields)
    for name, t, values in fields:
        if t == 1:
            s += f'<DataArray type="Float64" Name="{name}" format="ascii">\n' + _interior(values, n) + "</DataArray>\n"
        elif t == 2 and has_vectors:
            for d, suf in enumerate(("_x", "_y", "_z")):
                s += (f'<DataArray type="Float64" Name="{name}{suf}" format="ascii">\n'
                      + _interior(values, n, d * N) + "</DataArray>\n")
    s += "</CellData>\n</Piece>\n</ImageData>\n</VTKFile>\n"
    return s


def step_tag(step: int) -> str:
    return "%05d" % step


def writevth(n, step) -> str:
    tag = step_tag(step)
    return ('<?xml version="1.0"?>\n'
            '<VTKFile type="vtkOverlappingAMR" version="1.1" byte_order="LittleEndian" header_type="UInt32">\n'
            '<vtkOverlappingAMR origin="0 0 0" grid_description="XYZ">\n'
            '<Block level="0" spacing="1 1 1">\n'
            f'<DataSet index="0" amr_box="0 {n[0] - 2} 0 {n[1] - 2} 0 {n[2] - 2}" file="level/{tag}_level_0.vti">\n'
            '</DataSet>\n</Block>\n</vtkOverlappingAMR>\n</VTKFile>\n')
