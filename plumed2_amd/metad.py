"""METAD action on a GPU grid, host side.

Mirrors the keyword interface of the reference bias (src/bias/MetaD.cpp:
registerKeywords :640-721, grid set-up :940-1060, calculate :2315-2375,
update :2377-2520, HILLS records :1831-1884, readGaussians :1807-1829, grid
read at start-up :1450-1530) for diagonal, grid-based, optionally
well-tempered METAD, and the grid file format of src/tools/Grid.cpp
(writeHeader/writeToFile :523-580, GridBase::create from a file :621-700).  Deposition, the spline
bias and the forces run in the HIP library.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import shlex

import numpy as np

from . import _capi as A

KBOLTZMANN = 0.0083144621  # kJ/mol/K (src/tools/Tools.h:59)


class p2a_metad_spec(C.Structure):
    _fields_ = [("ncv", C.c_int32), ("periodic", C.c_int32 * 4), ("gmin", C.c_double * 4),
                ("gmax", C.c_double * 4), ("nbin", C.c_uint32 * 4), ("sigma", C.c_double * 4),
                ("height", C.c_double), ("biasf", C.c_double), ("kbt", C.c_double), ("pace", C.c_int32),
                ("nostretch", C.c_int32)]


_P = C.c_void_p
A.OPTIONAL_SIGS.update({
    "p2a_metad_create": ([C.POINTER(_P), C.c_int, C.POINTER(p2a_metad_spec)], C.c_int),
    "p2a_metad_destroy": ([_P], C.c_int),
    "p2a_metad_set_stream": ([_P, _P], C.c_int),
    "p2a_metad_grid_size": ([_P], C.c_int64),
    "p2a_metad_deposit": ([_P, C.c_int64, _P, _P], C.c_int),
    "p2a_metad_deposit_device": ([_P, C.c_int64, _P, _P], C.c_int),
    "p2a_metad_add_hills": ([_P, C.c_int64, _P, _P], C.c_int),
    "p2a_metad_bias": ([_P, C.c_int64, _P, _P, _P], C.c_int),
    "p2a_metad_bias_device": ([_P, C.c_int64, _P, _P, _P], C.c_int),
    "p2a_metad_step_device": ([_P, C.c_int64, _P, _P, _P, _P], C.c_int),
    "p2a_metad_check": ([_P], C.c_int),
    "p2a_metad_get_grid": ([_P, _P, _P], C.c_int),
    "p2a_metad_set_grid": ([_P, _P, _P], C.c_int),
    "p2a_apply_force_device": ([_P, C.c_int64, _P, _P, _P, _P, _P], C.c_int),
})
if A._lib is not None:  # library already loaded: bind the new symbols
    for _n, (_a, _r) in A.OPTIONAL_SIGS.items():
        _f = getattr(A._lib, _n)
        _f.argtypes, _f.restype = _a, _r


def _num(s: str) -> float:
    """Numbers as PLUMED accepts them in these keywords (incl. pi, -pi)."""
    t = s.strip().lower()
    sign = -1.0 if t.startswith("-") else 1.0
    t = t.lstrip("+-")
    if t == "pi":
        return sign * math.pi
    if t.endswith("*pi"):
        return sign * float(t[:-3]) * math.pi
    return sign * float(t)


def _words(line: str):
    return shlex.split(line.replace("{", " ").replace("}", " "))


class MetaD:
    """GPU METAD from a reference-syntax line, e.g.

        m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300
           GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512 FILE=HILLS FMT=%20.10f

    ``periodic`` maps an argument name to its (min, max) domain strings, as the
    upstream Value would declare it (e.g. {"phi": ("-pi", "pi")})."""

    KEYS = {"ARG", "SIGMA", "HEIGHT", "PACE", "BIASFACTOR", "TEMP", "GRID_MIN", "GRID_MAX", "GRID_BIN",
            "GRID_SPACING", "FILE", "FMT", "LABEL", "RESTART", "GRID_WFILE", "GRID_WSTRIDE", "GRID_RFILE"}

    def __init__(self, line: str, device: int = 0, periodic: dict | None = None, timestep: float = 1.0):
        words = _words(line)
        if words and words[0].endswith(":"):
            self.label = words.pop(0)[:-1]
        else:
            self.label = None
        if words and words[0] == "METAD":
            words.pop(0)
        kv, flags = {}, set()
        for w in words:
            if "=" in w:
                k, v = w.split("=", 1)
                kv[k] = v
            else:
                flags.add(w)
        unknown = [k for k in kv if k not in self.KEYS] + [f for f in flags if f not in {"GRID_SPARSE", "STORE_GRIDS"}]
        if unknown:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "cannot understand the following words from the input line : "
                                    + ", ".join(unknown))
        self.label = kv.get("LABEL", self.label)
        if "ARG" not in kv:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "keyword ARG is compulsory")
        self.args = kv["ARG"].split(",")
        ncv = len(self.args)
        if not 1 <= ncv <= 4:
            raise A.Plumed2AmdError(A.P2A_ERR_UNSUPPORTED, "METAD grid supports 1 to 4 CVs")
        sigma = [float(x) for x in kv.get("SIGMA", "").split(",") if x]
        if len(sigma) != ncv:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "number of arguments does not match number of SIGMA parameters")
        if "HEIGHT" not in kv:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "HEIGHT is compulsory (TAU is not implemented)")
        self.height0 = float(kv["HEIGHT"])
        self.pace = int(kv.get("PACE", "0") or 0)
        if self.pace <= 0:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "PACE should be positive")
        self.biasf = float(kv.get("BIASFACTOR", "-1.0"))
        temp = float(kv.get("TEMP", "-1"))
        self.kbt = KBOLTZMANN * temp if temp >= 0 else 0.0
        periodic = periodic or {}
        gmin_s = kv.get("GRID_MIN", "").split(",") if "GRID_MIN" in kv else []
        gmax_s = kv.get("GRID_MAX", "").split(",") if "GRID_MAX" in kv else []
        if len(gmin_s) != ncv or len(gmax_s) != ncv:
            raise A.Plumed2AmdError(A.P2A_ERR_UNSUPPORTED, "this METAD needs GRID_MIN and GRID_MAX for every ARG")
        gbin = [int(x) for x in kv["GRID_BIN"].split(",")] if "GRID_BIN" in kv else []
        if gbin and len(gbin) != ncv:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "not enough values for GRID_BIN")
        gspacing = [float(x) for x in kv["GRID_SPACING"].split(",")] if "GRID_SPACING" in kv else []
        if gspacing and len(gspacing) != ncv:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "not enough values for GRID_SPACING")
        if not gbin and not gspacing:
            gspacing = [0.2 * s for s in sigma]  # MetaD.cpp:960-967
        if not gbin:
            gbin = [1] * ncv
        gmin = [_num(x) for x in gmin_s]
        gmax = [_num(x) for x in gmax_s]
        if gspacing:  # MetaD.cpp:986-996: the larger number of bins wins
            for i in range(ncv):
                n = int(math.ceil((gmax[i] - gmin[i]) / gspacing[i]))
                gbin[i] = max(gbin[i], n)
        self.funcname = (self.label or "metad") + ".bias"
        self.wgrid_stride = int(kv.get("GRID_WSTRIDE", "0") or 0)
        self.wgrid_path = kv.get("GRID_WFILE")
        self.store_grids = "STORE_GRIDS" in flags
        if self.wgrid_path and self.wgrid_stride <= 0:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "frequency with which to output grid not specified use GRID_WSTRIDE")
        if self.wgrid_stride > 0 and not self.wgrid_path:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "grid filename not specified use GRID_WFILE")
        rgrid = None
        if kv.get("GRID_RFILE"):
            rpath = kv["GRID_RFILE"]
            if not os.path.exists(rpath):
                raise A.Plumed2AmdError(A.P2A_ERR_ARG, "The GRID file you want to read: " + rpath
                                        + ", cannot be found!")
            rgrid = read_grid_file(rpath, self.funcname, self.args)
            if len(rgrid["gmin"]) != ncv:
                raise A.Plumed2AmdError(A.P2A_ERR_ARG, "mismatch between dimensionality of input grid and number "
                                        "of arguments")
            for i, a in enumerate(self.args):
                if rgrid["periodic"][i] != (a in periodic):
                    raise A.Plumed2AmdError(A.P2A_ERR_ARG, "periodicity mismatch between arguments and input bias")
            # the grid read from the file replaces the one of GRID_MIN/MAX/BIN
            gmin_s, gmax_s = list(rgrid["gmin_s"]), list(rgrid["gmax_s"])
            gmin, gmax = list(rgrid["gmin"]), list(rgrid["gmax"])
            gbin = list(rgrid["nbin"])
        self.gmin_s, self.gmax_s = list(gmin_s), list(gmax_s)
        spec = p2a_metad_spec()
        spec.ncv = ncv
        self.periodic = []
        for i, a in enumerate(self.args):
            if a in periodic:
                lo, hi = (_num(x) for x in periodic[a])
                # a periodic argument's grid is its domain (Grid.cpp Init with pmin/pmax)
                gmin[i], gmax[i] = lo, hi
                self.gmin_s[i], self.gmax_s[i] = periodic[a]
                spec.periodic[i] = 1
                self.periodic.append(periodic[a])
            else:
                self.periodic.append(None)
            spec.gmin[i], spec.gmax[i], spec.nbin[i], spec.sigma[i] = gmin[i], gmax[i], gbin[i], sigma[i]
        spec.height, spec.biasf, spec.kbt, spec.pace = self.height0, self.biasf, self.kbt, self.pace
        spec.nostretch = 1 if os.environ.get("PLUMED_DP2CUTOFF_NOSTRETCH") else 0
        self.sigma = sigma
        self.gmin, self.gmax, self.gbin = gmin, gmax, gbin
        self.ncv = ncv
        self.welltemp = self.biasf >= 1.0
        self.timestep = timestep
        self.fmt = kv.get("FMT", "%f")
        self.hills_path = kv.get("FILE", "HILLS")
        self._hills = None
        self._ctx = C.c_void_p()
        A.check(A.lib().p2a_metad_create(C.byref(self._ctx), int(device), C.byref(spec)))
        self.size = A.lib().p2a_metad_grid_size(self._ctx)
        self._first = True
        self._append = False
        if rgrid is not None:
            self.set_grid(rgrid["values"], rgrid["derivs"])
        if kv.get("RESTART", "NO") == "YES" and os.path.exists(self.hills_path):
            self.read_hills(self.hills_path)
            self._append = True  # a restarted run appends to its HILLS file
        self._gridfile_open = False
        if self.wgrid_path:
            self._open_grid_file()

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            A.lib().p2a_metad_destroy(self._ctx)
            self._ctx = C.c_void_p()
        if self._hills:
            self._hills.close()
            self._hills = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- bias -----------------------------------------------------------
    def bias(self, cv):
        """Spline bias and gradient at one or more points (npoints x ncv)."""
        x = np.ascontiguousarray(np.asarray(cv, dtype=np.float64).reshape(-1, self.ncv))
        b = np.empty(len(x))
        d = np.empty((len(x), self.ncv))
        A.check(A.lib().p2a_metad_bias(self._ctx, len(x), x.ctypes.data, b.ctypes.data, d.ctypes.data))
        return b, d

    def calculate(self, cv):
        """MetaD::calculate: (bias, forces on the CVs = -dV/ds)."""
        if self.biasf == 1.0:
            return 0.0, np.zeros(self.ncv)
        b, d = self.bias(cv)
        return float(b[0]), -d[0]

    # ---- update -----------------------------------------------------------
    def update(self, step: int, cv):
        """MetaD::update: deposit a hill every PACE steps, not on the first
        step seen.  Returns the deposited height or None."""
        out = None
        if step % self.pace == 0 and not self._first:
            c = np.ascontiguousarray(np.asarray(cv, dtype=np.float64).reshape(1, self.ncv))
            h = np.empty(1)
            A.check(A.lib().p2a_metad_deposit(self._ctx, 1, c.ctypes.data, h.ctypes.data))
            self._write_hill(step * self.timestep, c[0], h[0])
            out = float(h[0])
        self._first = False
        # the grid is dumped after the step's hill, first step included
        if self.wgrid_stride > 0 and step % self.wgrid_stride == 0:
            self.write_grid()
        return out

    def deposit(self, centers):
        """Deposit many hills in sequence (one device launch); returns heights."""
        c = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, self.ncv))
        h = np.empty(len(c))
        A.check(A.lib().p2a_metad_deposit(self._ctx, len(c), c.ctypes.data, h.ctypes.data))
        return h

    def add_hills(self, centers, heights):
        c = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, self.ncv))
        h = np.ascontiguousarray(np.asarray(heights, dtype=np.float64).reshape(-1))
        A.check(A.lib().p2a_metad_add_hills(self._ctx, len(c), c.ctypes.data, h.ctypes.data))

    # ---- device step --------------------------------------------------------
    def step_device(self, step: int, cv_ptr: int, bias_ptr: int, force_ptr: int, height_ptr: int = 0,
                    stream: int | None = None):
        if stream is not None:
            A.check(A.lib().p2a_metad_set_stream(self._ctx, C.c_void_p(stream)))
        A.check(A.lib().p2a_metad_step_device(self._ctx, int(step), C.c_void_p(cv_ptr), C.c_void_p(bias_ptr),
                                              C.c_void_p(force_ptr), C.c_void_p(height_ptr)))

    def check(self):
        A.check(A.lib().p2a_metad_check(self._ctx))

    # ---- grid access ----------------------------------------------------------
    def grid(self):
        v = np.empty(self.size)
        d = np.empty((self.size, self.ncv))
        A.check(A.lib().p2a_metad_get_grid(self._ctx, v.ctypes.data, d.ctypes.data))
        return v, d

    def set_grid(self, values, derivs):
        v = np.ascontiguousarray(values, dtype=np.float64)
        d = np.ascontiguousarray(derivs, dtype=np.float64)
        A.check(A.lib().p2a_metad_set_grid(self._ctx, v.ctypes.data, d.ctypes.data))

    # ---- grid files ------------------------------------------------------------
    def node_counts(self):
        """Nodes per CV: a non-periodic grid has one more node than bins."""
        return [b if p is not None else b + 1 for b, p in zip(self.gbin, self.periodic)]

    def grid_text(self, fmt: str = "%14.9f"):
        """The grid as Grid::writeToFile prints it (Grid.cpp:523-580): the
        FIELDS line, min/max/nbins/periodic of every CV as constant fields
        (nbins counts nodes), then one line per node with the first CV
        running fastest and a blank line whenever the second-to-last index
        returns to zero."""
        v, d = self.grid()
        return format_grid(self.funcname, self.args, self.gmin_s, self.gmax_s, self.gmin, self.gmax, self.gbin,
                           [p is not None for p in self.periodic], v, d, fmt)

    def _open_grid_file(self):
        # OFile::open: an existing file is kept as bck.<n>.<name>
        _backup(self.wgrid_path)
        open(self.wgrid_path, "w").close()
        self._gridfile_open = True

    def write_grid(self, path: str | None = None):
        """Write the grid.  To GRID_WFILE (path None) as MetaD::update does
        (MetaD.cpp:2489-2520): the previous content is moved to
        bck.last.<name> and replaced (OFile::rewind), or, with STORE_GRIDS,
        the grid is appended with its header repeated."""
        text = self.grid_text()
        if path is not None:
            with open(path, "w") as f:
                f.write(text)
            return
        if not self._gridfile_open:
            self._open_grid_file()
        if self.store_grids:
            with open(self.wgrid_path, "a") as f:
                f.write(text)
            return
        head, tail = os.path.split(self.wgrid_path)
        os.replace(self.wgrid_path, os.path.join(head, "bck.last." + tail))
        with open(self.wgrid_path, "w") as f:
            f.write(text)

    # ---- HILLS files ------------------------------------------------------------
    def hills_header(self):
        names = " ".join(self.args)
        sig = " ".join("sigma_" + a for a in self.args)
        lines = [f"#! FIELDS time {names} {sig} height biasf", "#! SET multivariate false",
                 "#! SET kerneltype stretched-gaussian"]
        for a, p in zip(self.args, self.periodic):
            if p is not None:
                lines += [f"#! SET min_{a} {p[0]}", f"#! SET max_{a} {p[1]}"]
        return "\n".join(lines) + "\n"

    def hill_record(self, time, center, height):
        """MetaD::writeGaussian (MetaD.cpp:1831-1884): height stored as
        h*biasf/(biasf-1) for well-tempered runs."""
        h = height * self.biasf / (self.biasf - 1.0) if (self.welltemp and self.biasf > 1.0) else height
        fields = [time, *center, *self.sigma, h, self.biasf]
        return " ".join(self.fmt % f for f in fields) + "\n"

    def _write_hill(self, time, center, height):
        if self.hills_path is None:
            return
        if self._hills is None:
            self._hills = open(self.hills_path, "a" if self._append else "w")
            if not self._append:
                self._hills.write(self.hills_header())
        self._hills.write(self.hill_record(time, center, height))
        self._hills.flush()

    def read_hills(self, path):
        """MetaD::readGaussians: add the hills of a HILLS file to the grid
        (heights rescaled by (biasf-1)/biasf for well-tempered runs)."""
        names = None
        rows = []
        with open(path) as f:
            for line in f:
                if line.startswith("#! FIELDS"):
                    names = line.split()[2:]
                elif line.startswith("#") or not line.strip():
                    continue
                else:
                    rows.append([float(v) for v in line.split()])
        if not rows:
            return 0
        arr = np.array(rows)
        idx = [names.index(a) for a in self.args]
        h = arr[:, names.index("height")].copy()
        if self.welltemp and self.biasf > 1.0:
            h *= (self.biasf - 1.0) / self.biasf
        self.add_hills(arr[:, idx], h)
        return len(arr)


def _backup(path: str):
    """OFile's backup of an existing output file: renamed to the first free
    bck.<n>.<name> in the same directory."""
    if not os.path.exists(path):
        return
    head, tail = os.path.split(path)
    n = 0
    while os.path.exists(os.path.join(head, f"bck.{n}.{tail}")):
        n += 1
    os.replace(path, os.path.join(head, f"bck.{n}.{tail}"))


def format_grid(funcname, args, gmin_s, gmax_s, gmin, gmax, gbin, periodic, values, derivs, fmt="%14.9f"):
    """Text of a grid with derivatives in the layout of Grid::writeToFile
    (src/tools/Grid.cpp:523-580).  gbin are bins; the header's nbins counts
    nodes (bins + 1 when not periodic).  values[i] / derivs[i, :] belong to
    the flat node index i with the first CV running fastest."""
    ncv = len(args)
    nodes = [b if p else b + 1 for b, p in zip(gbin, periodic)]
    dx = [(hi - lo) / b for lo, hi, b in zip(gmin, gmax, gbin)]
    lines = ["#! FIELDS " + " ".join(list(args) + [funcname] + ["der_" + a for a in args])]
    for i, a in enumerate(args):
        lines += [f"#! SET min_{a} {gmin_s[i]}", f"#! SET max_{a} {gmax_s[i]}",
                  f"#! SET nbins_{a}  {nodes[i]}", f"#! SET periodic_{a} {'true' if periodic[i] else 'false'}"]
    f = " " + fmt
    total = int(np.prod(nodes))
    values = np.asarray(values).reshape(total)
    derivs = np.asarray(derivs).reshape(total, ncv)
    idx = [0] * ncv
    for k in range(total):
        r = k
        for j in range(ncv):
            idx[j] = r % nodes[j]
            r //= nodes[j]
        if k > 0 and ncv > 1 and idx[ncv - 2] == 0:
            lines.append("")
        x = [gmin[j] + idx[j] * dx[j] for j in range(ncv)]
        lines.append("".join(f % v for v in (*x, values[k], *derivs[k])))
    return "\n".join(lines) + "\n"


def read_grid_file(path, funcname, args):
    """Read a grid written as above (GridBase::create, Grid.cpp:621-700): the
    grid geometry comes from the header; each line's value and derivatives
    go to the node nearest its point + dx/2.  Returns a dict with gmin_s,
    gmax_s, gmin, gmax, nbin (bins), periodic, values, derivs."""
    fields, consts, rows = None, {}, []
    with open(path) as f:
        for line in f:
            if line.startswith("#! FIELDS"):
                fields = line.split()[2:]
            elif line.startswith("#! SET"):
                w = line.split()
                consts[w[2]] = w[3] if len(w) > 3 else ""
            elif line.startswith("#") or not line.strip():
                continue
            else:
                rows.append([float(v) for v in line.split()])
    if fields is None or funcname not in fields:
        raise A.Plumed2AmdError(A.P2A_ERR_ARG, "no column labelled " + funcname + " in in grid input")
    ncv = len(args)
    pos = []
    for a in args:
        if a not in fields:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, f"no column labelled {a} in grid input")
        pos.append(fields.index(a))
    if pos != sorted(pos):
        raise A.Plumed2AmdError(A.P2A_ERR_ARG, "arguments in input are not in same order as in grid file")
    out = {"gmin_s": [], "gmax_s": [], "gmin": [], "gmax": [], "nbin": [], "periodic": []}
    for a in args:
        try:
            lo_s, hi_s, nb, per = (consts["min_" + a], consts["max_" + a], int(consts["nbins_" + a]),
                                   consts["periodic_" + a])
        except KeyError as e:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, f"missing field {e.args[0]} in grid input") from None
        if nb <= 0:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "nbins must be positive in grid input")
        out["gmin_s"].append(lo_s)
        out["gmax_s"].append(hi_s)
        out["gmin"].append(_num(lo_s))
        out["gmax"].append(_num(hi_s))
        out["periodic"].append(per == "true")
        # the header counts nodes: one more than bins when not periodic
        out["nbin"].append(nb if per == "true" else nb - 1)
    dcols = []
    for a in args:
        if "der_" + a not in fields:
            raise A.Plumed2AmdError(A.P2A_ERR_ARG, "missing derivatives from grid file")
        dcols.append(fields.index("der_" + a))
    nodes = [b if p else b + 1 for b, p in zip(out["nbin"], out["periodic"])]
    total = int(np.prod(nodes))
    values = np.zeros(total)
    derivs = np.zeros((total, ncv))
    fcol = fields.index(funcname)
    for r in rows:
        k, stride = 0, 1
        for j in range(ncv):
            dx = (out["gmax"][j] - out["gmin"][j]) / out["nbin"][j]
            i = int(math.floor((r[pos[j]] + dx / 2.0 - out["gmin"][j]) / dx))
            i = i % nodes[j] if out["periodic"][j] else min(max(i, 0), nodes[j] - 1)
            k += i * stride
            stride *= nodes[j]
        values[k] = r[fcol]
        derivs[k] = [r[c] for c in dcols]
    out["values"], out["derivs"] = values, derivs
    return out
