// Checkpoint, visualisation and grid files in the reference's formats, so
// that restart files, the visualisation outputs and the post-processing
// readers of a reference run and of this library are interchangeable:
//   load_one / io_field        load.f90:84-187
//   out3d, write_log_output    output.f90:227-288, write_visu_3d 289-303
//   save_grid / load_grid      initgrid.f90:128-165
//   geometry_b_NNN.out         main.f90:432-437
// Arrays are the block's interior values in Fortran order (i fastest), FP64.
#pragma once
#include <string>
#include <vector>

namespace snac {

// Fortran Ew.d edit descriptor (0.ddd...E+ee, right-justified in w columns)
std::string fortran_e(double x, int w, int d);
// "NNN", the 1-based block number written with (i3.3)
std::string block_num(int b);

// load_one('w'): the n values, then [time, istep] (load.f90:164-175)
void write_field_file(const std::string &path, const double *v, size_t n, double time, int istep);
// load_one('r'): the file must hold exactly n + 2 doubles (load.f90:110-121,
// the reference aborts otherwise); reads the n values, time and nint(istep)
void read_field_file(const std::string &path, double *v, size_t n, double &time, int &istep);
// out3d: the n values only (output.f90:250-256)
void write_raw_field(const std::string &path, const double *v, size_t n);
// write_log_output: appends one line '(A,A,A,9i5,E15.7,i10)' (output.f90:281-285)
void append_visu_log(const std::string &path, const std::string &fname, const std::string &varname,
                     const int nmin[3], const int nmax[3], const int nskip[3], double time, int istep);
// save_grid: <base>.bin holds rf, rc, drf, drc over lo..hi; <base>.out holds
// lo-1..hi+1 as '(5E15.7)' rows 0, rf, rc, drf, drc.  Arrays have n+2
// entries, index 0 = lo-1.
void save_grid(const std::string &base, int n, const std::vector<double> &rf, const std::vector<double> &rc,
               const std::vector<double> &drf, const std::vector<double> &drc);
// load_grid: when <base>.bin exists, reads rf, rc, drf, drc over lo..hi and
// rebuilds the ghost entries as initgrid.f90:155-164 does; false when absent
bool load_grid(const std::string &base, int n, std::vector<double> &rf, std::vector<double> &rc,
               std::vector<double> &drf, std::vector<double> &drc);
// geometry_b_NNN.out: lo_g, hi_g, lmin, lmax, one list-directed row each
void write_geometry(const std::string &path, const int lo[3], const int hi[3], const double lmin[3],
                    const double lmax[3]);

}  // namespace snac
