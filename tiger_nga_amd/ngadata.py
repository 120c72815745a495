"""NGA data (restart) files: read, write, and the chemistry fields they hold.

The file layout is that of the reference's parallel reader and writer
(core/data.f90:218-284 ``data_read``, 379-457 ``data_write_full3D``) and of
the serial writer of its case generators (tools/init_flow/param.f90:113-148
``param_write_data``), all in native byte order:

    int32   nx, ny, nz, nvars                      16 bytes
    float64 dt, time                               16 bytes
    nvars names of str_short = 8 characters,       8 * nvars bytes
      blank padded (library/string.f90:3)
    nvars arrays of nx*ny*nz float64 in Fortran order (i fastest);
      variable v (1-based) starts at byte
      16 + 8*nvars + 16 + nx*ny*nz*8*(v-1)         (data.f90:271-272)

Variables named rhoU rhoV rhoW U V W P RHO dRHO VISC DIFF are flow fields;
every other variable is a transported scalar, numbered in file order
(data.f90:109-154).  The chemistry step reads RHO and the scalars: the
species, whose names must follow the mechanism's species order from the
first species on (combustion/finitechem.f90:218-251), and T
(combustion/combustion.f90:393-409).

Reading maps each variable with ``numpy.memmap`` (no copy until a variable
is used); ``chemistry_block`` assembles the Fortran-layout RHO_ / SC_ block
(ghost layers included) that ``FiniteChem.source_chemistry`` takes, on the
host or directly in device memory.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

STR_SHORT = 8  # library/string.f90:3
HEADER_INTS = 4
# names that data_mpi_init links to flow fields rather than to SC
# (core/data.f90:109-137); every other name is a scalar
FLOW_NAMES = ("rhoU", "rhoV", "rhoW", "U", "V", "W", "P", "RHO", "dRHO", "VISC", "DIFF")

_I4 = np.dtype("=i4")
_F8 = np.dtype("=f8")


class DataFileError(ValueError):
    """A data file that does not match the grid or the mechanism (the
    reference stops the run with ``die`` in these cases)."""


def header_bytes(nvars: int) -> int:
    """Bytes before the first variable: 4*4 + str_short*nvars + 2*8."""
    return 4 * HEADER_INTS + STR_SHORT * nvars + 2 * _F8.itemsize


def var_offset(nx: int, ny: int, nz: int, nvars: int, var: int) -> int:
    """Byte offset of variable ``var`` (1-based), core/data.f90:271-272."""
    if not 1 <= var <= nvars:
        raise IndexError("variable %d outside 1..%d" % (var, nvars))
    return header_bytes(nvars) + int(nx) * int(ny) * int(nz) * _F8.itemsize * (var - 1)


def _pack_name(name: str) -> bytes:
    """A Fortran character(len=8) assignment: truncate, then blank-pad."""
    b = name.encode("ascii")[:STR_SHORT]
    return b + b" " * (STR_SHORT - len(b))


def _unpack_name(raw: bytes) -> str:
    """trim(): trailing blanks removed, as every reference comparison does."""
    return raw.decode("ascii").rstrip(" ")


@dataclass
class DataFile:
    """One data file, header parsed and variables memory-mapped read-only."""
    path: str
    nx: int
    ny: int
    nz: int
    dt: float
    time: float
    names: list = field(default_factory=list)

    @property
    def nvars(self) -> int:
        return len(self.names)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def sc_names(self):
        """Scalar names in file order: SC(..., isc) is the isc-th of these
        (core/data.f90:148-154)."""
        return [n for n in self.names if n not in FLOW_NAMES]

    def scalar_index(self, name: str) -> int:
        """1-based scalar index of ``name``, 0 when absent
        (combustion/combustion.f90:393-409)."""
        sc = self.sc_names
        return sc.index(name) + 1 if name in sc else 0

    def var(self, name_or_index, copy_on_write: bool = False) -> np.ndarray:
        """Variable by name or 1-based file index, as a read-only (nx,ny,nz)
        Fortran-ordered memory map (copy_on_write: a writable private map
        whose changes never reach the file)."""
        if isinstance(name_or_index, str):
            if name_or_index not in self.names:
                raise KeyError("variable %r not in %s" % (name_or_index, self.path))
            v = self.names.index(name_or_index) + 1
        else:
            v = int(name_or_index)
        off = var_offset(self.nx, self.ny, self.nz, self.nvars, v)
        return np.memmap(self.path, dtype=_F8, mode="c" if copy_on_write else "r", offset=off, shape=self.shape,
                         order="F")


def read(path: str, expect_shape=None) -> DataFile:
    """Header of a data file (core/data.f90:237-258).  ``expect_shape`` is the
    grid's (nx, ny, nz): a different size raises, as data_read stops with
    'The size of the data file does not correspond to the grid file'."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        head = f.read(4 * HEADER_INTS + 2 * _F8.itemsize)
        if len(head) < 32:
            raise DataFileError("%s: truncated header" % path)
        nx, ny, nz, nvars = (int(x) for x in np.frombuffer(head, _I4, HEADER_INTS))
        dt, time = (float(x) for x in np.frombuffer(head, _F8, 2, offset=4 * HEADER_INTS))
        if min(nx, ny, nz) < 1 or nvars < 0:
            raise DataFileError("%s: bad dimensions %d %d %d, nvars %d" % (path, nx, ny, nz, nvars))
        raw = f.read(STR_SHORT * nvars)
    if len(raw) != STR_SHORT * nvars:
        raise DataFileError("%s: truncated variable names" % path)
    names = [_unpack_name(raw[STR_SHORT * v:STR_SHORT * (v + 1)]) for v in range(nvars)]
    need = header_bytes(nvars) + nx * ny * nz * _F8.itemsize * nvars
    if size < need:
        raise DataFileError("%s: %d bytes, the header announces %d" % (path, size, need))
    if expect_shape is not None and tuple(int(s) for s in expect_shape) != (nx, ny, nz):
        raise DataFileError("The size of the data file does not correspond to the grid file: grid = %s, "
                            "data = %s" % (tuple(expect_shape), (nx, ny, nz)))
    return DataFile(path, nx, ny, nz, dt, time, names)


def _host(a) -> np.ndarray:
    if isinstance(a, np.ndarray):
        return a
    return a.detach().cpu().numpy()  # torch tensor (host or device)


def write(path: str, variables, dt: float = 0.0, time: float = 0.0) -> None:
    """Write a data file (core/data.f90:412-446): ``variables`` is a sequence
    of (name, array) or a dict in file order; every array is (nx, ny, nz)
    (numpy or torch, any layout, written in Fortran order)."""
    items = list(variables.items()) if isinstance(variables, dict) else list(variables)
    if not items:
        raise ValueError("no variables to write")
    shape = tuple(_host(items[0][1]).shape)
    if len(shape) != 3:
        raise ValueError("variables must be 3-D (nx, ny, nz)")
    for name, a in items:
        if tuple(_host(a).shape) != shape:
            raise ValueError("variable %r has shape %s, expected %s" % (name, _host(a).shape, shape))
    with open(path, "wb") as f:
        f.write(np.array(shape + (len(items),), dtype=_I4).tobytes())
        f.write(np.array([dt, time], dtype=_F8).tobytes())
        for name, _ in items:
            f.write(_pack_name(name))
        for _, a in items:
            np.ravel(np.asarray(_host(a), dtype=_F8), order="F").tofile(f)


def timestamped_name(filename: str, time: float) -> str:
    """File name used when 'Data overwrite' is 0: the name, '_', and the time
    in ES12.3 with blanks removed (core/data.f90:399-403); e.g. 1.5e-3 gives
    ``name_1.500E-03``."""
    m, e = ("%.3E" % time).split("E")
    ex = int(e)
    # ES12.3 writes a two-digit exponent as E+dd and a three-digit one as +ddd
    es = "%sE%s%02d" % (m, "-" if ex < 0 else "+", abs(ex)) if abs(ex) < 100 else "%s%+04d" % (m, ex)
    return filename.strip() + "_" + es


def species_offset(sc_names, species):
    """(isc_sc, isc) of the chemistry's species in the scalar list:
    isc_sc is the 1-based scalar index of the mechanism's first species, the
    following scalars must be the mechanism's species in order, names compared
    on their first str_short characters (combustion/finitechem.f90:218-251).
    isc maps H2O, CO2, CH4 and CO to their 1-based scalar index (0: absent)."""
    first = species[0][:STR_SHORT]
    if first not in sc_names:
        raise DataFileError("Species %s not initialized. Run init flow again" % species[0])
    isc_sc = list(sc_names).index(first) + 1
    isc = {"H2O": 0, "CO2": 0, "CH4": 0, "CO": 0}
    for k, sp in enumerate(species):
        i = isc_sc + k
        if i > len(sc_names) or sc_names[i - 1] != sp[:STR_SHORT]:
            raise DataFileError("Species %s not found in initial file" % sp)
        if sc_names[i - 1] in isc:
            isc[sc_names[i - 1]] = i
    return isc_sc, isc


def chemistry_block(df: DataFile, species, nover: int = 0, device=None):
    """RHO_, SC_ and the Grid of a single block holding the whole file, with
    ``nover`` ghost layers: Fortran-ordered float64, RHO_ (nx+2n, ny+2n,
    nz+2n), SC_ (..., nscalar).  Interior cells come from the file; ghost
    cells keep data_init's defaults (RHO 1, SC 0; core/data.f90:322-343) and a
    file without RHO leaves the interior at 1 as well.  ``device`` (e.g.
    'cuda') places both arrays in device memory; they are then filled by one
    host-to-device copy per variable, read straight from the memory map."""
    from .finitechem import Grid
    sc = df.sc_names
    isc_sc, _ = species_offset(sc, species)
    isc_T = df.scalar_index("T")
    if isc_T == 0:
        raise DataFileError("%s: no temperature scalar 'T'" % df.path)
    g = Grid.block(df.nx, df.ny, df.nz, len(sc), isc_sc, isc_T, nover=nover)
    sh = g.shape
    core = (slice(nover, nover + df.nx), slice(nover, nover + df.ny), slice(nover, nover + df.nz))
    if device is None:
        RHO_ = np.ones(sh, order="F")
        SC_ = np.zeros(sh + (len(sc),), order="F")
        if "RHO" in df.names:
            RHO_[core] = df.var("RHO")
        for s, name in enumerate(sc):
            SC_[core + (s,)] = df.var(name)
        return RHO_, SC_, g
    import torch

    def fortran(shape):
        st, acc = [], 1
        for s in shape:
            st.append(acc)
            acc *= s
        return torch.empty_strided(shape, tuple(st), dtype=torch.float64, device=device)
    RHO_ = fortran(sh).fill_(1.0)
    SC_ = fortran(sh + (len(sc),)).zero_()
    if "RHO" in df.names:
        RHO_[core].copy_(torch.from_numpy(np.asarray(df.var("RHO", copy_on_write=True))))
    for s, name in enumerate(sc):
        SC_[core + (s,)].copy_(torch.from_numpy(np.asarray(df.var(name, copy_on_write=True))))
    return RHO_, SC_, g


def write_chemistry_block(path: str, df: DataFile, RHO_, SC_, grid, dt=None, time=None) -> None:
    """Write ``df``'s variables back with RHO and the scalars taken from the
    block's interior (flow variables other than RHO copied from ``df``), in
    ``df``'s variable order -- the data_write_full3D of a run whose state
    lives in RHO_ / SC_."""
    o = (grid.imin_ - grid.imino_, grid.jmin_ - grid.jmino_, grid.kmin_ - grid.kmino_)
    core = tuple(slice(o[d], o[d] + n) for d, n in enumerate(df.shape))
    sc = df.sc_names
    items = []
    for name in df.names:
        if name == "RHO":
            items.append((name, _host(RHO_[core])))
        elif name in FLOW_NAMES:
            items.append((name, df.var(name)))
        else:
            items.append((name, _host(SC_[core + (sc.index(name),)])))
    write(path, items, df.dt if dt is None else dt, df.time if time is None else time)
