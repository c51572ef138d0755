// Host-side description of a SNaC multi-block case: inputs (param.f90), block
// placement and neighbours (initmpi.f90), grids (initgrid.f90), inflow planes
// and pressure BC rhs planes.  Pure C++, no device code.
#pragma once
#include <array>
#include <map>
#include <string>
#include <vector>

namespace snac {

// One value of a namelist entry: number, string or logical.
struct NmlValue {
  enum Kind { Num, Str, Bool } kind = Num;
  double num = 0.0;
  std::string str;
  bool b = false;
};

// group -> variable -> (index tuple -> value); scalars use the empty tuple.
using NmlGroup = std::map<std::string, std::map<std::vector<int>, NmlValue>>;
using Nml = std::map<std::string, NmlGroup>;

// Fortran namelist reader for the subset SNaC's inputs use (see nml.cpp).
Nml read_namelist_file(const std::string &path);

struct Block {
  std::array<int, 3> n{}, lo_g{};
  std::array<double, 3> lmin{}, lmax{}, gr{};
  std::array<int, 3> gt{};
  char cbcvel[3][3][2];   // [comp][dir][side]
  double bcvel[3][3][2];
  char cbcpre[3][2];      // [dir][side]
  double bcpre[3][2];
  int inflow_type[3][2];
  std::string inivel = "zer";  // block_inivel, the initial velocity field
  int nb[3][2];           // neighbour block per [dir][side], -1 = boundary
  bool is_bound[3][2];
  bool is_inflow[3][2];
  // metrics, n+2 entries (index 0 = lo-1): dc = d?c, df = d?f, rc/rf coords
  std::array<std::vector<double>, 3> dc, df, rc, rf;
  // df, dc as initgrid (or a grid file) left them, before bound_grid
  // overwrites the interface ghosts: what save_grid writes
  std::array<std::vector<double>, 3> dc0, df0;
  // inflow planes [dir][comp][side] over the ghosted tangential plane
  std::vector<double> vin[3][3][2];
  // pressure BC rhs planes [dir][side] over the interior tangential plane
  std::vector<double> rhsp[3][2];
  // scalar BCs, flat [iscal][dir][side] (block_cbcscal / block_bcscal)
  std::vector<char> cbcsca;
  std::vector<double> bcsca;
};

struct Case {
  // &dns
  double cfl = 0, dtmax = 0, dt_f = -1, visci = 1, visc = 1, time_max = 0, tw_max = 0;
  int nstep = 0, icheck = 0, nscal = 0, nsaves_max = 0;
  int iout0d = 0, iout1d = 0, iout2d = 0, iout3d = 0, isave = 0;
  bool stop_type[3] = {false, false, false}, restart = false, is_overwrite_save = true;
  double bforce[3] = {0, 0, 0}, gacc[3] = {0, 0, 0};
  // &hypre
  int hypre_solver_i = 2, hypre_maxiter = 50;
  double hypre_tol = 1e-4;
  // &scalar (param.f90:96-100,148-163); alpha = 1/alphai
  std::vector<double> alpha, ssource, scalf;
  std::vector<int> is_sforced;
  std::vector<std::string> iniscal;
  double beta = 0.0, alpha_max = 0.0;
  // derived
  int fft_axis = 0;  // 0 none, 1..3
  std::vector<Block> blocks;
  int periods[3] = {0, 0, 0};
  bool is_periodic[3] = {false, false, false};
  double vol_all = 0;
};

// Builds a case from the two input files; throws std::runtime_error with the
// reference's error wording on invalid input.
Case build_case(const std::string &dns_nml, const std::string &blocks_nml, int fft_axis);

// initgrid.f90:9-83 for one direction (outputs n+2 entries each).
void initgrid(int n, int gt, double gr, double lmin, double lmax, std::vector<double> &drc,
              std::vector<double> &drf, std::vector<double> &rc, std::vector<double> &rf);

// fft.f90:96-149 and 151-189: centred BC pairs (the pressure, scalars and
// the velocity components across their own axis) and, with centred false,
// the staggered pairs FF and DD (a velocity component along its own axis;
// DD transforms the first n-1 points, its n-th eigenvalue is left zero)
bool find_fft(char bc0, char bc1, int &kind_fwd, int &kind_bwd, double norm[2], bool centred = true);
bool eigenvalues(int n, char bc0, char bc1, std::vector<double> &lambda, bool centred = true);

}  // namespace snac
