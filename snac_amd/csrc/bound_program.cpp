// Ghost-fill programs.  bounduvw/boundp/inflow (bound.f90:8-106, 438-477)
// are expanded into a sequence of plane operations in the reference's order
// (halos per direction x, y, z -- the per-field order of the reference is
// immaterial because fields do not interact -- then set_bc per block in the
// order x0, x1, y0, y1, z0, z1, then the inflow planes).  Consecutive
// operations are packed into one phase as long as no operation of the phase
// writes a plane another one reads or writes; each phase is one kernel
// launch.  Two planes of one array intersect when their normals differ
// (they share an edge line) or when they are the same plane.  This keeps the
// edge and corner ghost values exactly as the reference's sequential fill
// leaves them, which matters because the advection stencil reads edge
// ghosts (e.g. v(i+1,j-1,k) in momx_a_vv, mom.f90:112-147).
#include "bound_program.hpp"

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace snac {
namespace {

struct Rec {
  FaceOp op;
  const void *warr, *rarr;  // array identities (null when no read)
  int stage = 0;            // 0..2 halo x/y/z, 3 set_bc, 4 inflow
};

bool planes_meet(const void *a1, int d1, int q1, const void *a2, int d2, int q2) {
  if (!a1 || !a2 || a1 != a2) return false;
  return d1 != d2 || q1 == q2;
}

bool conflict(const Rec &x, const Rec &y) {
  if (planes_meet(x.warr, x.op.d, x.op.qdst, y.warr, y.op.d, y.op.qdst)) return true;
  if (planes_meet(x.warr, x.op.d, x.op.qdst, y.rarr, y.op.d, y.op.qsrc)) return true;
  if (planes_meet(y.warr, y.op.d, y.op.qdst, x.rarr, x.op.d, x.op.qsrc)) return true;
  return false;
}

void tang(int d, int &a1, int &a2) {
  a1 = d == 0 ? 1 : 0;
  a2 = d == 2 ? 1 : 2;
}

Rec mk(const BlockArrays &B, double *dst, int d, int qdst, const double *src, const BlockArrays *SB, int qsrc,
       const double *arr, double c0, double c1, double c2) {
  Rec r;
  FaceOp &o = r.op;
  int a1, a2;
  tang(d, a1, a2);
  o.dst = dst;
  o.src = src;
  o.arr = arr;
  o.dsx = B.g.sx;
  o.dsy = B.g.sy;
  const BlockArrays &S = SB ? *SB : B;
  o.ssx = S.g.sx;
  o.ssy = S.g.sy;
  o.d = d;
  o.qdst = qdst;
  o.qsrc = qsrc;
  int n3[3] = {B.g.nx, B.g.ny, B.g.nz};
  o.m1 = n3[a1] + 2;
  o.m2 = n3[a2] + 2;
  o.c0 = c0;
  o.c1 = c1;
  o.c2 = c2;
  r.warr = dst;
  r.rarr = src;
  return r;
}

// set_bc bound.f90:108-248 as plane operations
void set_bc_ops(std::vector<Rec> &out, const BlockArrays &B, double *p, char ctype, int side, int d, bool centered,
                double rvalue, double dr) {
  if (ctype != 'D' && ctype != 'N') return;
  int n3[3] = {B.g.nx, B.g.ny, B.g.nz};
  int n = n3[d];
  double factor = rvalue, sgn = 0.0;
  if (ctype == 'D' && centered) { factor = 2.0 * factor; sgn = -1.0; }
  if (ctype == 'N') { factor = (side == 0 ? -dr : dr) * factor; sgn = 1.0; }
  if (centered) {
    if (side == 0) out.push_back(mk(B, p, d, 0, p, nullptr, 1, nullptr, factor, sgn, 0.0));
    else out.push_back(mk(B, p, d, n + 1, p, nullptr, n, nullptr, factor, sgn, 0.0));
  } else if (ctype == 'D') {
    if (side == 0) out.push_back(mk(B, p, d, 0, nullptr, nullptr, 0, nullptr, factor, 0.0, 0.0));
    else {
      out.push_back(mk(B, p, d, n, nullptr, nullptr, 0, nullptr, factor, 0.0, 0.0));
      out.push_back(mk(B, p, d, n + 1, p, nullptr, n - 1, nullptr, 0.0, 1.0, 0.0));
    }
  } else {
    if (side == 0) out.push_back(mk(B, p, d, 0, p, nullptr, 1, nullptr, factor, 1.0, 0.0));
    else {
      out.push_back(mk(B, p, d, n, p, nullptr, n - 1, nullptr, factor, 1.0, 0.0));
      out.push_back(mk(B, p, d, n + 1, p, nullptr, n - 1, nullptr, factor, 1.0, 0.0));
    }
  }
}

void halo_ops(std::vector<Rec> &out, const std::vector<BlockArrays> &bl, const Case &cs, int d,
              double *(*field)(const BlockArrays &, int), int fid) {
  size_t first = out.size();
  for (size_t b = 0; b < bl.size(); b++) {
    const Block &B = cs.blocks[b];
    int n3[3] = {B.n[0], B.n[1], B.n[2]};
    for (int side = 0; side < 2; side++) {
      int nb = B.nb[d][side];
      if (nb < 0) continue;
      int qdst = side == 0 ? 0 : n3[d] + 1;
      int qsrc = side == 0 ? cs.blocks[nb].n[d] : 1;
      out.push_back(mk(bl[b], field(bl[b], fid), d, qdst, field(bl[nb], fid), &bl[nb], qsrc, nullptr, 0.0, 1.0, 0.0));
    }
  }
  for (size_t q = first; q < out.size(); q++) out[q].stage = d;
}

}  // namespace

namespace {
std::vector<std::vector<FaceOp>> pack(const std::vector<Rec> &seq) {
  std::vector<std::vector<FaceOp>> phases;
  std::vector<Rec> cur;
  for (const Rec &r : seq) {
    bool clash = false;
    for (const Rec &c : cur)
      if (conflict(r, c)) { clash = true; break; }
    if (clash) {
      std::vector<FaceOp> ph;
      for (auto &c : cur) ph.push_back(c.op);
      phases.push_back(ph);
      cur.clear();
    }
    cur.push_back(r);
  }
  if (!cur.empty()) {
    std::vector<FaceOp> ph;
    for (auto &c : cur) ph.push_back(c.op);
    phases.push_back(ph);
  }
  return phases;
}
}  // namespace

namespace {
void tag_from(std::vector<Rec> &seq, size_t first, int stage) {
  for (size_t q = first; q < seq.size(); q++) seq[q].stage = stage;
}

std::vector<Rec> velocity_ops(const Case &cs, const std::vector<BlockArrays> &bl, int par, bool is_correc,
                              bool with_halo_bc, bool with_inflow) {
  std::vector<Rec> seq;
  std::vector<BlockArrays> sel = bl;
  for (auto &B : sel) B.par_sel = par;
  auto field = [](const BlockArrays &B, int comp) -> double * { return B.vel[B.par_sel][comp]; };
  if (with_halo_bc) {
    for (int d = 0; d < 3; d++)
      for (int comp = 0; comp < 3; comp++) halo_ops(seq, sel, cs, d, field, comp);
    size_t first_bc = seq.size();
    for (size_t b = 0; b < sel.size(); b++) {
      const Block &B = cs.blocks[b];
      for (int d = 0; d < 3; d++)
        for (int side = 0; side < 2; side++) {
          if (!B.is_bound[d][side]) continue;
          int qm = side == 0 ? 0 : B.n[d];
          for (int comp = 0; comp < 3; comp++) {
            bool centered = comp != d;
            if (!centered && is_correc) continue;
            double dr = centered ? B.dc[d][qm] : B.df[d][qm];
            set_bc_ops(seq, sel[b], field(sel[b], comp), B.cbcvel[comp][d][side], side, d, centered,
                       B.bcvel[comp][d][side], dr);
          }
        }
    }
    tag_from(seq, first_bc, 3);
  }
  if (with_inflow) {
    size_t first_in = seq.size();
    for (size_t b = 0; b < sel.size(); b++) {
      const Block &B = cs.blocks[b];
      for (int d = 0; d < 3; d++)
        for (int side = 0; side < 2; side++) {
          if (!B.is_inflow[d][side]) continue;
          int n = B.n[d];
          int qn = side == 0 ? 0 : n, qt = side == 0 ? 0 : n + 1, dq = side == 0 ? 1 : -1;
          for (int comp = 0; comp < 3; comp++) {
            double *F = field(sel[b], comp);
            const double *arr = sel[b].vin[d][comp][side];
            if (comp == d) {
              if (!is_correc) seq.push_back(mk(sel[b], F, d, qn, nullptr, nullptr, 0, arr, 0.0, 0.0, 1.0));
            } else {
              seq.push_back(mk(sel[b], F, d, qt, F, nullptr, qt + dq, arr, 0.0, -1.0, 2.0));
            }
          }
        }
    }
    tag_from(seq, first_in, 4);
  }
  return seq;
}

std::vector<Rec> pressure_ops(const Case &cs, const std::vector<BlockArrays> &bl, int which) {
  std::vector<Rec> seq;
  std::vector<BlockArrays> sel = bl;
  for (auto &B : sel) B.par_sel = which;  // 0 = p, 1 = pp
  auto field = [](const BlockArrays &B, int) -> double * { return B.par_sel == 0 ? B.p : B.pp; };
  for (int d = 0; d < 3; d++) halo_ops(seq, sel, cs, d, field, 0);
  size_t first_bc = seq.size();
  for (size_t b = 0; b < sel.size(); b++) {
    const Block &B = cs.blocks[b];
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B.is_bound[d][side]) continue;
        int qm = side == 0 ? 0 : B.n[d];
        set_bc_ops(seq, sel[b], field(sel[b], 0), B.cbcpre[d][side], side, d, true, B.bcpre[d][side], B.dc[d][qm]);
      }
  }
  tag_from(seq, first_bc, 3);
  return seq;
}

// boundp with the scalar's BCs (B.cbcsca / B.bcsca), same rules as pressure
std::vector<Rec> scalar_ops(const Case &cs, const std::vector<BlockArrays> &bl, int par, int iscal) {
  std::vector<Rec> seq;
  std::vector<BlockArrays> sel = bl;
  for (auto &B : sel) B.par_sel = par;
  auto field = [](const BlockArrays &B, int is) -> double * { return B.sv[B.par_sel][is]; };
  for (int d = 0; d < 3; d++) halo_ops(seq, sel, cs, d, field, iscal);
  size_t first_bc = seq.size();
  for (size_t b = 0; b < sel.size(); b++) {
    const Block &B = cs.blocks[b];
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        if (!B.is_bound[d][side]) continue;
        int qm = side == 0 ? 0 : B.n[d];
        const size_t k = (size_t)iscal * 6 + d * 2 + side;
        set_bc_ops(seq, sel[b], field(sel[b], iscal), B.cbcsca[k], side, d, true, B.bcsca[k], B.dc[d][qm]);
      }
  }
  tag_from(seq, first_bc, 3);
  return seq;
}

// Fields do not interact, so the ops of several fills may be interleaved as
// long as each array keeps its own order: a stable sort by stage does that.
std::vector<std::vector<FaceOp>> merge_and_pack(std::vector<Rec> seq) {
  std::stable_sort(seq.begin(), seq.end(), [](const Rec &a, const Rec &b) { return a.stage < b.stage; });
  return pack(seq);
}
}  // namespace

std::vector<std::vector<FaceOp>> build_velocity_program(const Case &cs, const std::vector<BlockArrays> &bl, int par,
                                                        bool is_correc, bool with_halo_bc, bool with_inflow) {
  return merge_and_pack(velocity_ops(cs, bl, par, is_correc, with_halo_bc, with_inflow));
}

std::vector<std::vector<FaceOp>> build_pressure_program(const Case &cs, const std::vector<BlockArrays> &bl, int which) {
  return merge_and_pack(pressure_ops(cs, bl, which));
}

std::vector<std::vector<FaceOp>> build_scalar_program(const Case &cs, const std::vector<BlockArrays> &bl, int par,
                                                      int iscal) {
  std::vector<Rec> seq;
  for (int is = 0; is < cs.nscal; is++) {
    if (iscal >= 0 && is != iscal) continue;
    std::vector<Rec> one = scalar_ops(cs, bl, par, is);
    seq.insert(seq.end(), one.begin(), one.end());
  }
  return merge_and_pack(seq);
}

std::vector<std::vector<FaceOp>> build_correc_fill_program(const Case &cs, const std::vector<BlockArrays> &bl,
                                                           int par) {
  std::vector<Rec> seq = velocity_ops(cs, bl, par, true, true, true);
  std::vector<Rec> ps = pressure_ops(cs, bl, 0);
  seq.insert(seq.end(), ps.begin(), ps.end());
  return merge_and_pack(seq);
}

namespace {

long plane_index(int d, int q, int t1, int t2, long sx, long sy) {
  if (d == 0) return (long)q + sx * ((long)t1 + sy * (long)t2);
  if (d == 1) return (long)t1 + sx * ((long)q + sy * (long)t2);
  return (long)t1 + sx * ((long)t2 + sy * (long)q);
}

struct CellKey {
  const void *a;
  long i;
  bool operator==(const CellKey &o) const { return a == o.a && i == o.i; }
};
struct CellKeyHash {
  size_t operator()(const CellKey &k) const {
    return std::hash<const void *>()(k.a) ^ (std::hash<long>()(k.i) * 0x9E3779B97F4A7C15ull);
  }
};

// the chain of one written cell, with the op (and plane row) that wrote it last
struct Chain {
  const double *base;
  long bidx;
  int nlev;
  CellLev lev[SEG_LEV];
  double *dst;
  long didx;
  int op, t1, t2;
};

bool same_levels(const Chain &a, const Chain &b) {
  if (a.nlev != b.nlev || (a.base == nullptr) != (b.base == nullptr) || a.base != b.base) return false;
  for (int l = 0; l < a.nlev; l++) {
    const CellLev &x = a.lev[l], &y = b.lev[l];
    if (std::memcmp(&x.c0, &y.c0, sizeof(double)) || std::memcmp(&x.c1, &y.c1, sizeof(double)) ||
        std::memcmp(&x.c2, &y.c2, sizeof(double)) || x.arr != y.arr || x.src != y.src)
      return false;
  }
  return true;
}

}  // namespace

bool compose_program(const std::vector<std::vector<FaceOp>> &phases, std::vector<CellSeg> &segs) {
  std::unordered_map<CellKey, int, CellKeyHash> at;
  std::vector<Chain> ch;
  size_t ncell = 0;
  for (auto &ph : phases)
    for (auto &o : ph) ncell += (size_t)o.m1 * o.m2;
  at.reserve(ncell);
  ch.reserve(ncell);
  int opi = 0;
  for (auto &ph : phases)
    for (auto &o : ph) {
      opi++;
      for (int t2 = 0; t2 < o.m2; t2++)
        for (int t1 = 0; t1 < o.m1; t1++) {
          Chain e;
          const long t = t1 + (long)o.m1 * t2;
          const CellLev lv{o.c0, o.c1, o.c2, o.arr, o.arr ? t : 0, o.src ? 1 : 0};
          if (o.src) {
            const long sidx = plane_index(o.d, o.qsrc, t1, t2, o.ssx, o.ssy);
            auto it = at.find({o.src, sidx});
            if (it != at.end()) {
              e = ch[it->second];
            } else {
              e.base = o.src;
              e.bidx = sidx;
              e.nlev = 0;
            }
            if (e.nlev == SEG_LEV) return false;
            e.lev[e.nlev++] = lv;
          } else {
            e.base = nullptr;
            e.bidx = 0;
            e.nlev = 1;
            e.lev[0] = lv;
          }
          e.dst = o.dst;
          e.didx = plane_index(o.d, o.qdst, t1, t2, o.dsx, o.dsy);
          e.op = opi;
          e.t1 = t1;
          e.t2 = t2;
          auto it = at.find({o.dst, e.didx});
          if (it == at.end()) {
            at.emplace(CellKey{o.dst, e.didx}, (int)ch.size());
            ch.push_back(e);
          } else {
            ch[it->second] = e;
          }
        }
    }
  // a final value that read a cell the program writes (after the read) would
  // race with that write in one launch
  for (const Chain &e : ch)
    if (e.base && at.count({e.base, e.bidx})) return false;
  std::vector<int> ord(ch.size());
  for (size_t q = 0; q < ch.size(); q++) ord[q] = (int)q;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) {
    const Chain &a = ch[x], &b = ch[y];
    if (a.op != b.op) return a.op < b.op;
    if (a.t2 != b.t2) return a.t2 < b.t2;
    return a.t1 < b.t1;
  });
  std::vector<CellSeg> out;
  size_t q = 0;
  while (q < ord.size()) {
    const Chain &f = ch[ord[q]];
    size_t r = q + 1;
    long dstr = 0, bstr = 0;
    while (r < ord.size()) {
      const Chain &prev = ch[ord[r - 1]], &e = ch[ord[r]];
      if (e.op != f.op || e.t2 != f.t2 || e.t1 != prev.t1 + 1 || e.dst != f.dst || !same_levels(f, e)) break;
      const long ds = e.didx - prev.didx, bs = e.base ? e.bidx - prev.bidx : 0;
      if (r == q + 1) {
        dstr = ds;
        bstr = bs;
      } else if (ds != dstr || bs != bstr) {
        break;
      }
      bool arr_ok = true;
      for (int l = 0; l < e.nlev; l++)
        if (e.lev[l].arr && e.lev[l].a0 != prev.lev[l].a0 + 1) arr_ok = false;
      if (!arr_ok) break;
      r++;
    }
    CellSeg g;
    std::memset(&g, 0, sizeof g);
    g.dst = f.dst;
    g.d0 = f.didx;
    g.dstr = dstr;
    g.base = f.base;
    g.b0 = f.bidx;
    g.bstr = bstr;
    g.len = (int)(r - q);
    g.nlev = f.nlev;
    for (int l = 0; l < f.nlev; l++) g.lev[l] = f.lev[l];
    out.push_back(g);
    q = r;
  }
  segs.swap(out);
  return true;
}

}  // namespace snac
