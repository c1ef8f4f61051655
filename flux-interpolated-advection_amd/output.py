"""The reference's output files, written from the host side of this library:

* density dumps, reference src/write_substance_density_data.cpp:11-27:
  <path>/substance_density_<N>.dat = "nx ny nz" then one value per line in [x][y][z] order (z
  fastest), in iostream's default format (printf "%g", 6 significant digits);
* the per-frame time series, reference src/main.cpp:159-210 (files) and 518-597 (values):
  <path>/<prefix><name>.dat, one "time value" line per frame at 16 significant digits, for total
  mass, total momenta, the momentum the walls took out through the pressure, and their sums.

Also a binary density dump (int32 nx, ny, nz, then float32 values in the same order), which this
library adds for restarts: it keeps every bit, the text dump keeps 6 digits.
"""
import os

import numpy as np

SERIES = ["total_mass", "total_momentum_x", "total_momentum_y", "total_momentum_z", "total_loss_of_momentum_x",
          "total_loss_of_momentum_y", "total_loss_of_momentum_z", "momentum_x_plus_total_loss_of_momentum_x",
          "momentum_y_plus_total_loss_of_momentum_y", "momentum_z_plus_total_loss_of_momentum_z"]


def _host(a):
    if isinstance(a, np.ndarray):
        return a
    return a.detach().cpu().numpy()


def write_substance_density_data(density, file_number, density_data_path):
    """Text dump of a [nx, ny, nz] density field (numpy array or device tensor)."""
    d = np.ascontiguousarray(_host(density), dtype=np.float32)
    nx, ny, nz = d.shape
    name = os.path.join(density_data_path, "substance_density_%d.dat" % file_number)
    with open(name, "w") as f:
        f.write("%d %d %d\n" % (nx, ny, nz))
        f.write("".join("%g\n" % v for v in d.reshape(-1).tolist()))
    return name


def read_substance_density_data(filename):
    with open(filename) as f:
        nx, ny, nz = (int(t) for t in f.readline().split())
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != nx * ny * nz:
        raise ValueError("%s: %d values for a %dx%dx%d grid" % (filename, vals.size, nx, ny, nz))
    return vals.astype(np.float32).reshape(nx, ny, nz)


def write_substance_density_raw(density, filename):
    d = np.ascontiguousarray(_host(density), dtype=np.float32)
    with open(filename, "wb") as f:
        np.asarray(d.shape, dtype="<i4").tofile(f)
        d.astype("<f4").tofile(f)


def read_substance_density_raw(filename):
    with open(filename, "rb") as f:
        shape = tuple(int(v) for v in np.fromfile(f, dtype="<i4", count=3))
        d = np.fromfile(f, dtype="<f4")
    if d.size != shape[0] * shape[1] * shape[2]:
        raise ValueError("%s: truncated" % filename)
    return d.reshape(shape)


class SeriesWriter:
    """The ten *.dat series of the reference; `write` appends one frame.  The files are truncated
    when the writer opens them, as the reference removes and recreates them at start-up."""

    def __init__(self, path, prefix=""):
        os.makedirs(path, exist_ok=True)
        self.files = {k: open(os.path.join(path, prefix + k + ".dat"), "w") for k in SERIES}

    def write(self, now_time, totals, loss):
        """totals = (mass, momentum x, y, z); loss = accumulated wall loss (x, y, z)."""
        mass, mx, my, mz = totals
        lx, ly, lz = loss
        vals = [mass, mx, my, mz, lx, ly, lz, mx + lx, my + ly, mz + lz]
        for k, v in zip(SERIES, vals):
            self.files[k].write("%.16g %.16g\n" % (now_time, v))
            self.files[k].flush()

    def close(self):
        for f in self.files.values():
            f.close()


def read_series(filename):
    """[[time, value], ...] of one series file."""
    return np.loadtxt(filename, ndmin=2)
