"""File formats of SNaC's checkpoint, visualisation and grid files,
restated in numpy: the checker of the library's writers and readers.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).  Paths
cited are relative to the SNaC tree.

  load_one / io_field   src/load.f90:84-187   interior values in Fortran
                        order (the block's ng(1)*ng(2)*ng(3) doubles), then
                        [time, istep] as two doubles
  out3d                 src/output.f90:227-257 the interior values only
  write_log_output      src/output.f90:259-288 '(A,A,A,9i5,E15.7,i10)'
  save_grid             src/initgrid.f90:128-143 <f>.bin: rf, rc, drf, drc
                        over lo..hi (one direct-access record); <f>.out:
                        rows lo-1..hi+1 of '(5E15.7)' 0, rf, rc, drf, drc
  load_grid ghosts      src/initgrid.f90:155-164

fortran_e follows the Fortran standard's Ew.d edit descriptor
(0.d1d2...dd E+ee, rounded to nearest); its pins are worked examples of that
rule in tests/test_fileformats.py.  The readers of the reference
(utils/read_binary_data/python/read_single_field_binary.py,
read_restart_file.py) take the same layouts: reshape(..., order='F').
"""
import numpy as np


def field_file_bytes(interior, time, istep):
    """load_one('w') of one field (load.f90:164-175)."""
    a = np.asarray(interior, dtype=np.float64)
    return a.reshape(-1, order="F").tobytes() + np.array([time, float(istep)]).tobytes()


def read_field_file(raw, shape):
    """load_one('r') (load.f90:103-140): (interior, time, istep); None when
    the size is not that of the shape plus two doubles."""
    v = np.frombuffer(raw, dtype=np.float64)
    n = int(np.prod(shape))
    if v.size != n + 2:
        return None
    return v[:n].reshape(shape, order="F"), float(v[n]), int(np.rint(v[n + 1]))


def fortran_e(x, w=15, d=7):
    """Ew.d: sign, '0.', d digits of the mantissa in [0.1, 1), 'E' and a
    signed two-digit exponent (three digits without the 'E' past 99)."""
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
        body = ("-" if np.signbit(x) else "") + body
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        e = int(e) + 1
        ex = f"E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) <= 99 else f"{'-' if e < 0 else '+'}{abs(e):03d}"
        body = ("-" if x < 0 else "") + "0." + digits + ex
    return body.rjust(w) if len(body) <= w else "*" * w


def visu_log_line(fname, varname, nmin, nmax, nskip, time, istep):
    """One line of write_log_output (output.f90:281-285)."""
    ints = "".join(f"{int(i):5d}" for i in list(nmin) + list(nmax) + list(nskip))
    return f"{fname} {varname}{ints}{fortran_e(time)}{int(istep):10d}"


def grid_bin_bytes(rf, rc, drf, drc):
    """save_grid's .bin record: the lo..hi entries of each array (arrays with
    ghosts, index 0 = lo-1)."""
    return b"".join(np.asarray(a[1:-1], dtype=np.float64).tobytes() for a in (rf, rc, drf, drc))


def grid_out_lines(rf, rc, drf, drc):
    """save_grid's .out rows lo-1..hi+1."""
    return ["".join(fortran_e(v) for v in (0.0, rf[q], rc[q], drf[q], drc[q])) for q in range(len(rf))]


def load_grid_ghosts(rf, rc, drf, drc):
    """The ghost entries load_grid rebuilds from the lo..hi values
    (initgrid.f90:155-164); arrays with ghosts, index 0 = lo-1."""
    rf, rc, drf, drc = (np.array(a, dtype=np.float64) for a in (rf, rc, drf, drc))
    n = len(rf) - 2
    drf[0] = drf[1]
    drf[n + 1] = drf[n]
    drc[0] = drf[1]
    drc[n + 1] = drc[n]
    rf[0] = rf[1] - drf[1]
    rf[n + 1] = rf[n] + drf[n]
    rc[0] = rc[1] - drc[0]
    rc[n + 1] = rc[n] + drc[n]
    return rf, rc, drf, drc
