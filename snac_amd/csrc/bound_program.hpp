// Ghost-fill programs: per-phase lists of FaceOps (see bound_program.cpp).
#pragma once
#include <vector>

#include "case.hpp"
#include "common.hpp"

namespace snac {

// Device arrays of one block that the fill programs address.
struct BlockArrays {
  Geo g;
  double *vel[2][3];          // velocity ping-pong buffers [parity][u,v,w]
  double *p, *pp;
  std::vector<double *> sv[2];  // scalar ping-pong buffers [parity][iscal]
  const double *vin[3][3][2]; // device inflow planes [dir][comp][side] (null if none)
  int par_sel = 0;            // selector used while building
};

// bounduvw (halo + set_bc, when with_halo_bc) and/or inflow on the velocity
// buffers of parity `par`; returns the phases in execution order.
std::vector<std::vector<FaceOp>> build_velocity_program(const Case &cs, const std::vector<BlockArrays> &bl, int par,
                                                        bool is_correc, bool with_halo_bc, bool with_inflow);
// boundp on p (which = 0) or pp (which = 1)
std::vector<std::vector<FaceOp>> build_pressure_program(const Case &cs, const std::vector<BlockArrays> &bl, int which);
// boundp of scalar iscal (or of every scalar when iscal < 0) in the buffers
// of parity `par`, with the scalars' own BCs (main.f90:565-568,934)
std::vector<std::vector<FaceOp>> build_scalar_program(const Case &cs, const std::vector<BlockArrays> &bl, int par,
                                                      int iscal);
// bounduvw(is_correc) + inflow(is_correc) + boundp(p), the three fills that end
// a substage (main.f90:1026-1032), merged into shared phases
std::vector<std::vector<FaceOp>> build_correc_fill_program(const Case &cs, const std::vector<BlockArrays> &bl,
                                                           int par);
// Composes a phased program into cell segments that one launch can run:
// every written cell's chain of ops is followed back to a cell the program
// does not write.  False (segs untouched) when a final value depends on a
// cell the program writes later (it read the old value there) or a chain is
// longer than SEG_LEV; the phased program must run then.
bool compose_program(const std::vector<std::vector<FaceOp>> &phases, std::vector<CellSeg> &segs);

}  // namespace snac
