// Problem set-up on the host: master mesh, domain meshes and the RWG edge
// numbering without periodic boundaries, and the classification of edges
// on mirror planes by the mesh's line elements (the `sbnd` command).
//
// Reference algorithm (for90-mom2): load_mesh_gmsh (mesh.f90:610-736),
// build_mesh (mesh.f90:1044-1137),
// compute_basis_data (1265-1322), average_edge_length (138-153),
// extract_submesh (290-376), submesh_edge_connectivity (378-452),
// orient_basis (454-558); the `mesh` and `sdom` commands that drive them
// (interface.f90:357-366, 44-105).
//
// The reference finds edges and parents with nested scans (O(F E) and
// O(E^2)); here hash maps keyed by the sorted node pair give the same
// first-appearance numbering in one pass.  Every rule that fixes an index,
// a sign or a rounding is kept:
//  - edges are numbered by first appearance in face order, the pair being
//    nodes (s, s+1) of the face; the first face holding an edge is its "+"
//    face and that face's third vertex the "+" free vertex; a later face is
//    the "-" face (with more than two, the last one);
//  - geometry uses the reference's operation order with FMA contraction
//    off, so the arrays carry the reference's bits;
//  - a domain holds the master faces with a listed id, in master order, its
//    nodes the used master nodes in ascending order, the nodes of faces with
//    a negative id reversed;
//  - each domain edge takes the master edge with the same node pair as its
//    parent; the children of a master edge are oriented against the first
//    child, and, as in the reference, a success found only in the last pass
//    is reported as a failure and ends the orientation.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

struct HostMesh {
  std::vector<double> p;                // [V][3]
  std::vector<int32_t> fnodes, fid;     // [F][3], [F]
  std::vector<int32_t> fparent, nparent;  // master face of each face, master node of each node
  std::vector<int32_t> fedges;          // [F][3]
  std::vector<int32_t> enodes, ebn, efaces;  // [E][2]
  std::vector<int32_t> eparent;         // [E]
  std::vector<double> fn, fcp, fs, fm, farea, elen;
  double avelen = 0;
  int64_t V() const { return (int64_t)p.size() / 3; }
  int64_t F() const { return (int64_t)fid.size(); }
  int64_t E() const { return (int64_t)eparent.size(); }
};

static inline uint64_t pair_key(int32_t a, int32_t b) {
  const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// build_mesh (mesh.f90:1044-1137): first-appearance edge numbering
static void build_edges(HostMesh &m) {
  const int64_t F = m.F();
  std::unordered_map<uint64_t, int32_t> lookup;
  lookup.reserve((size_t)(2 * F));
  m.fedges.assign(3 * F, -1);
  m.enodes.clear();
  m.ebn.clear();
  m.efaces.clear();
  for (int64_t f = 0; f < F; f++) {
    const int32_t *t = &m.fnodes[3 * f];
    for (int s = 0; s < 3; s++) {
      const int32_t a = t[s], b = t[(s + 1) % 3], c = t[(s + 2) % 3];
      auto it = lookup.find(pair_key(a, b));
      int32_t l;
      if (it == lookup.end()) {
        l = (int32_t)(m.enodes.size() / 2);
        lookup.emplace(pair_key(a, b), l);
        m.enodes.push_back(a);
        m.enodes.push_back(b);
        m.ebn.push_back(c);
        m.ebn.push_back(-1);
        m.efaces.push_back((int32_t)f);
        m.efaces.push_back(-1);
      } else {
        l = it->second;
        m.efaces[2 * l + 1] = (int32_t)f;  // more than two faces: the last is kept (mesh.f90:1082-1086)
        m.ebn[2 * l + 1] = c;
      }
      m.fedges[3 * f + s] = l;
    }
  }
  m.eparent.assign(m.enodes.size() / 2, -1);
}

// SQRT(SUM(v*v)) in the order v1*v1 + v2*v2 + v3*v3 (linalg.f90:54-59)
static inline double norm3(const double *v) {
#pragma clang fp contract(off)
  return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

static inline void cross3(const double *a, const double *b, double *r) {
#pragma clang fp contract(off)
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// compute_basis_data (mesh.f90:1265-1322) and average_edge_length (138-153)
static void basis_data(HostMesh &m) {
#pragma clang fp contract(off)
  const int64_t F = m.F(), E = m.E();
  m.fn.assign(3 * F, 0);
  m.fcp.assign(3 * F, 0);
  m.fs.assign(9 * F, 0);
  m.fm.assign(9 * F, 0);
  m.farea.assign(F, 0);
  for (int64_t f = 0; f < F; f++) {
    const double *p1 = &m.p[3 * m.fnodes[3 * f]], *p2 = &m.p[3 * m.fnodes[3 * f + 1]],
                 *p3 = &m.p[3 * m.fnodes[3 * f + 2]];
    double d21[3], d31[3], d32[3], d13[3], b[3];
    for (int c = 0; c < 3; c++) {
      d21[c] = p2[c] - p1[c];
      d31[c] = p3[c] - p1[c];
      d32[c] = p3[c] - p2[c];
      d13[c] = p1[c] - p3[c];
    }
    cross3(d21, d31, b);
    const double nb = norm3(b);
    double *n = &m.fn[3 * f];
    for (int c = 0; c < 3; c++) n[c] = b[c] / nb;
    const double *ed[3] = {d21, d32, d13};
    for (int e = 0; e < 3; e++) {
      const double ne = norm3(ed[e]);
      double *s = &m.fs[9 * f + 3 * e];
      for (int c = 0; c < 3; c++) s[c] = ed[e][c] / ne;
      double mm[3];
      cross3(s, n, mm);
      const double nm = norm3(mm);
      double *mo = &m.fm[9 * f + 3 * e];
      for (int c = 0; c < 3; c++) mo[c] = mm[c] / nm;
    }
    double m13[3], a[3];
    for (int c = 0; c < 3; c++) m13[c] = -d13[c];
    cross3(m13, d32, a);
    m.farea[f] = 0.5 * norm3(a);
    for (int c = 0; c < 3; c++) m.fcp[3 * f + c] = (p1[c] + p2[c] + p3[c]) / 3.0;
  }
  m.elen.assign(E, 0);
  double acc = 0;
  for (int64_t e = 0; e < E; e++) {
    const double *a = &m.p[3 * m.enodes[2 * e]], *b = &m.p[3 * m.enodes[2 * e + 1]];
    const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    m.elen[e] = norm3(d);
    acc = acc + m.elen[e];
  }
  m.avelen = E > 0 ? acc / (double)E : 0.0;
}

static void build_mesh(HostMesh &m, double scale) {
#pragma clang fp contract(off)
  for (double &x : m.p) x = x * scale;
  build_edges(m);
  basis_data(m);
}

// extract_submesh (mesh.f90:290-376), surface faces only
static HostMesh extract(const HostMesh &master, const std::vector<int32_t> &sids) {
  HostMesh d;
  std::vector<char> want_face(master.F(), 0), used(master.V(), 0);
  for (int64_t f = 0; f < master.F(); f++) {
    for (int32_t s : sids)
      if (master.fid[f] == std::abs(s)) want_face[f] = 1;
    if (want_face[f])
      for (int c = 0; c < 3; c++) used[master.fnodes[3 * f + c]] = 1;
  }
  std::vector<int32_t> loc(master.V(), -1);
  for (int64_t v = 0; v < master.V(); v++)
    if (used[v]) {
      loc[v] = (int32_t)d.nparent.size();
      d.nparent.push_back((int32_t)v);
      for (int c = 0; c < 3; c++) d.p.push_back(master.p[3 * v + c]);
    }
  for (int64_t f = 0; f < master.F(); f++)
    if (want_face[f]) {
      for (int c = 0; c < 3; c++) d.fnodes.push_back(loc[master.fnodes[3 * f + c]]);
      d.fid.push_back(master.fid[f]);
      d.fparent.push_back((int32_t)f);
    }
  return d;
}

// submesh_edge_connectivity (mesh.f90:378-452), then orient_basis
// (mesh.f90:454-558)
static int connect_and_orient(const HostMesh &master, std::vector<HostMesh> &doms) {
  std::unordered_map<uint64_t, int32_t> lookup;
  lookup.reserve((size_t)(2 * master.E()));
  for (int64_t e = 0; e < master.E(); e++) lookup.emplace(pair_key(master.enodes[2 * e], master.enodes[2 * e + 1]), (int32_t)e);
  std::vector<std::vector<std::pair<int, int32_t>>> children(master.E());
  for (int d = 0; d < (int)doms.size(); d++) {
    HostMesh &dm = doms[d];
    for (int64_t l = 0; l < dm.E(); l++) {
      const int32_t a = dm.nparent[dm.enodes[2 * l]], b = dm.nparent[dm.enodes[2 * l + 1]];
      auto it = lookup.find(pair_key(a, b));
      if (it != lookup.end()) {
        dm.eparent[l] = it->second;
        children[it->second].push_back({d, (int32_t)l});
      }
    }
  }
  for (const auto &ch : children)
    if (ch.size() < 2) return fail(MOM_ERR_ARG, "Submesh decomposition is inconsistent (invalid child edges)!");
  for (const auto &ch : children) {
    const int nchild = (int)ch.size();
    std::vector<char> oriented(nchild, 0);
    oriented[0] = 1;
    int exited_at = -1;
    for (int it = 1; it <= nchild; it++) {
      for (int l = 1; l < nchild; l++) {
        if (oriented[l]) continue;
        HostMesh &Ml = doms[ch[l].first];
        const int32_t el = ch[l].second;
        for (int k = 0; k < nchild; k++) {
          if (!oriented[k]) continue;
          const HostMesh &Mk = doms[ch[k].first];
          const int32_t ek = ch[k].second;
          int32_t pf1 = Ml.efaces[2 * el], nf1 = Ml.efaces[2 * el + 1];
          int32_t pf2 = Mk.efaces[2 * ek], nf2 = Mk.efaces[2 * ek + 1];
          pf1 = pf1 != -1 ? Ml.fparent[pf1] : -1;
          nf1 = nf1 != -1 ? Ml.fparent[nf1] : -1;
          pf2 = pf2 != -1 ? Mk.fparent[pf2] : -1;
          nf2 = nf2 != -1 ? Mk.fparent[nf2] : -1;
          if (pf1 == pf2 || nf1 == nf2) {
            std::swap(Ml.efaces[2 * el], Ml.efaces[2 * el + 1]);
            std::swap(Ml.ebn[2 * el], Ml.ebn[2 * el + 1]);
          }
          if (pf1 == pf2 || nf1 == nf2 || pf1 == nf2 || nf1 == pf2) {
            oriented[l] = 1;
            break;
          }
        }
      }
      if (std::all_of(oriented.begin(), oriented.end(), [](char c) { return c != 0; })) {
        exited_at = it;
        break;
      }
    }
    if (exited_at == nchild) {
      std::fprintf(stderr, "Basis orientation was unsuccessful!\n");
      break;
    }
  }
  return 0;
}

// load_mesh_gmsh (mesh.f90:610-736): Gmsh 2.1/2.2 ASCII; node coordinates
// stored at their node number (numbers 1..nnodes), triangles (element type
// 2) with their first tag as the surface id; other elements are skipped.
static std::string trimmed(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

static bool parse_longs(const std::string &line, std::vector<long> &out) {
  out.clear();
  const char *c = line.c_str();
  while (true) {
    while (*c == ' ' || *c == '\t' || *c == '\r') c++;
    if (!*c) return true;
    char *end;
    errno = 0;
    const long v = std::strtol(c, &end, 10);
    if (end == c || errno) return false;
    out.push_back(v);
    c = end;
  }
}

static int read_gmsh(const char *path, std::vector<double> &p, std::vector<int32_t> &fnodes,
                     std::vector<int32_t> &fids, std::vector<int32_t> &lnodes, std::vector<int32_t> &lids) {
  std::ifstream in(path);
  if (!in) return fail(MOM_ERR_ARG, std::string("Could not open mesh file! ") + path);
  std::string line;
  bool have_nodes = false, have_elems = false;
  std::vector<long> v;
  auto expect_end = [&](const char *tag, const char *msg) -> int {
    if (!std::getline(in, line) || trimmed(line) != tag) return fail(MOM_ERR_ARG, msg);
    return 0;
  };
  while (std::getline(in, line)) {
    const std::string tok = trimmed(line);
    if (tok == "$MeshFormat") {
      if (!std::getline(in, line)) return fail(MOM_ERR_ARG, "Mesh version is unsupported!");
      const std::string ver = trimmed(line).substr(0, 3);
      if (ver != "2.1" && ver != "2.2") return fail(MOM_ERR_ARG, "Mesh version is unsupported!");
      MOM_TRY(expect_end("$EndMeshFormat", "Could not find $MeshFormatNodes specifier!"));
    } else if (tok == "$Nodes") {
      if (!std::getline(in, line) || !parse_longs(line, v) || v.size() != 1 || v[0] < 1 || v[0] > (1L << 28))
        return fail(MOM_ERR_ARG, "bad node count in mesh file");
      const long nn = v[0];
      p.assign(3 * nn, 0.0);
      std::vector<char> seen(nn, 0);
      for (long n = 0; n < nn; n++) {
        if (!std::getline(in, line)) return fail(MOM_ERR_ARG, "mesh file ends inside $Nodes");
        const char *c = line.c_str();
        char *end;
        const long num = std::strtol(c, &end, 10);
        if (end == c || num < 1 || num > nn) return fail(MOM_ERR_ARG, "node number out of range in mesh file");
        double x[3];
        for (int k = 0; k < 3; k++) {
          c = end;
          x[k] = std::strtod(c, &end);
          if (end == c) return fail(MOM_ERR_ARG, "bad node line in mesh file");
        }
        for (int k = 0; k < 3; k++) p[3 * (num - 1) + k] = x[k];
        seen[num - 1] = 1;
      }
      if (std::count(seen.begin(), seen.end(), 0)) return fail(MOM_ERR_ARG, "node numbers are not 1..nnodes");
      MOM_TRY(expect_end("$EndNodes", "Could not find $EndNodes specifier!"));
      have_nodes = true;
    } else if (tok == "$Elements") {
      if (!std::getline(in, line) || !parse_longs(line, v) || v.size() != 1 || v[0] < 0)
        return fail(MOM_ERR_ARG, "bad element count in mesh file");
      const long ne = v[0];
      fnodes.clear();
      fids.clear();
      lnodes.clear();
      lids.clear();
      for (long n = 0; n < ne; n++) {
        if (!std::getline(in, line)) return fail(MOM_ERR_ARG, "mesh file ends inside $Elements");
        if (!parse_longs(line, v) || v.size() < 3) return fail(MOM_ERR_ARG, "bad element line in mesh file");
        if (v[1] == 1) {
          // line element: its first tag is the id that `sbnd` refers to (mesh.f90:692-697)
          const long nt = v[2];
          if (nt < 1 || (long)v.size() < 5 + nt) return fail(MOM_ERR_ARG, "bad line element in mesh file");
          lids.push_back((int32_t)v[3]);
          for (int k = 0; k < 2; k++) lnodes.push_back((int32_t)(v[3 + nt + k] - 1));
          continue;
        }
        if (v[1] != 2) continue;
        const long nt = v[2];
        if (nt < 1 || (long)v.size() < 6 + nt) return fail(MOM_ERR_ARG, "bad triangle line in mesh file");
        fids.push_back((int32_t)v[3]);
        for (int k = 0; k < 3; k++) fnodes.push_back((int32_t)(v[3 + nt + k] - 1));
      }
      MOM_TRY(expect_end("$EndElements", "Could not find $EndElements specifier!"));
      have_elems = true;
    }
  }
  if (!have_nodes || !have_elems || fids.empty()) return fail(MOM_ERR_ARG, "mesh file has no nodes or no triangles");
  const long nn = (long)(p.size() / 3);
  for (int32_t i : fnodes)
    if (i < 0 || i >= nn) return fail(MOM_ERR_ARG, "face node index out of range");
  for (int32_t i : lnodes)
    if (i < 0 || i >= nn) return fail(MOM_ERR_ARG, "line node index out of range");
  return 0;
}

}  // namespace mom

using namespace mom;

struct mom_problem {
  HostMesh master;
  std::vector<HostMesh> doms;
  std::vector<int32_t> lnodes, lids;  // line elements (node pairs, ids) for edge classification
  std::vector<int32_t> bnd;           // boundary class of each master edge (mesh_bnd_*, 1 = none)
};

extern "C" {

int mom_problem_build(int64_t nnodes, const double *p, int64_t nfaces, const int32_t *fnodes, const int32_t *fids,
                      double scale, int32_t ndom, const int32_t *dom_ptr, const int32_t *dom_ids,
                      mom_problem **prob) {
  if (!p || !fnodes || !fids || !dom_ptr || !dom_ids || !prob) return fail(MOM_ERR_ARG, "NULL argument");
  if (nnodes < 3 || nfaces < 1 || ndom < 1 || nfaces > (1 << 28) || nnodes > (1 << 28))
    return fail(MOM_ERR_ARG, "mesh sizes out of range");
  if (!(scale > 0)) return fail(MOM_ERR_ARG, "scale must be > 0");
  for (int64_t i = 0; i < 3 * nfaces; i++)
    if (fnodes[i] < 0 || fnodes[i] >= nnodes) return fail(MOM_ERR_ARG, "face node index out of range");
  if (dom_ptr[0] != 0) return fail(MOM_ERR_ARG, "dom_ptr[0] must be 0");
  for (int d = 0; d < ndom; d++)
    if (dom_ptr[d + 1] <= dom_ptr[d]) return fail(MOM_ERR_ARG, "every domain needs at least one surface id");
  mom_problem *P = new mom_problem;
  P->master.p.assign(p, p + 3 * nnodes);
  P->master.fnodes.assign(fnodes, fnodes + 3 * nfaces);
  P->master.fid.assign(fids, fids + nfaces);
  P->master.fparent.resize(nfaces);
  for (int64_t f = 0; f < nfaces; f++) P->master.fparent[f] = (int32_t)f;
  P->master.nparent.resize(nnodes);
  for (int64_t v = 0; v < nnodes; v++) P->master.nparent[v] = (int32_t)v;
  build_mesh(P->master, scale);
  for (int d = 0; d < ndom; d++) {
    std::vector<int32_t> sids(dom_ids + dom_ptr[d], dom_ids + dom_ptr[d + 1]);
    for (int32_t s : sids)
      if (s == 0) {
        delete P;
        return fail(MOM_ERR_ARG, "Zero is not a valid surface id!");
      }
    HostMesh dm = extract(P->master, sids);
    for (int32_t s : sids)
      if (s < 0)
        for (int64_t f = 0; f < dm.F(); f++)
          if (dm.fid[f] == -s) std::swap(dm.fnodes[3 * f], dm.fnodes[3 * f + 2]);
    build_mesh(dm, 1.0);
    P->doms.push_back(std::move(dm));
  }
  const int rc = connect_and_orient(P->master, P->doms);
  if (rc) {
    delete P;
    return rc;
  }
  *prob = P;
  return 0;
}

int mom_problem_load_gmsh(const char *path, double scale, int32_t ndom, const int32_t *dom_ptr,
                          const int32_t *dom_ids, mom_problem **prob) {
  if (!path || !prob) return fail(MOM_ERR_ARG, "NULL argument");
  std::vector<double> p;
  std::vector<int32_t> fnodes, fids, lnodes, lids;
  MOM_TRY(read_gmsh(path, p, fnodes, fids, lnodes, lids));
  MOM_TRY(mom_problem_build((int64_t)(p.size() / 3), p.data(), (int64_t)fids.size(), fnodes.data(), fids.data(),
                            scale, ndom, dom_ptr, dom_ids, prob));
  (*prob)->lnodes = std::move(lnodes);
  (*prob)->lids = std::move(lids);
  return 0;
}

int mom_problem_set_lines(mom_problem *prob, int64_t nlines, const int32_t *nodes, const int32_t *ids) {
  if (!prob || (nlines > 0 && (!nodes || !ids)) || nlines < 0) return fail(MOM_ERR_ARG, "NULL argument");
  const int64_t V = prob->master.V();
  for (int64_t i = 0; i < 2 * nlines; i++)
    if (nodes[i] < 0 || nodes[i] >= V) return fail(MOM_ERR_ARG, "line node index out of range");
  prob->lnodes.assign(nodes, nodes + 2 * nlines);
  prob->lids.assign(ids, ids + nlines);
  return 0;
}

// classify_edges (mesh.f90:155-174) as the `sbnd` command runs it on the
// master mesh (interface.f90:120-157): every edge whose node pair is a line
// of the given id gets the boundary class; the domain edges inherit their
// parent's class (extract_submesh's inheritance, mesh.f90:427-445).
int mom_problem_classify_edges(mom_problem *prob, int32_t id, int32_t bnd, int64_t *nclassified) {
  if (!prob) return fail(MOM_ERR_ARG, "NULL argument");
  if (bnd < 1 || bnd > 10) return fail(MOM_ERR_ARG, "boundary class must be one of mesh_bnd_* (1..10)");
  HostMesh &M = prob->master;
  if (prob->bnd.empty()) prob->bnd.assign(M.E(), kBndNone);
  std::unordered_map<uint64_t, int64_t> count;
  for (size_t l = 0; l < prob->lids.size(); l++)
    if (prob->lids[l] == id) count[pair_key(prob->lnodes[2 * l], prob->lnodes[2 * l + 1])]++;
  int64_t n = 0;
  for (int64_t e = 0; e < M.E(); e++) {
    auto it = count.find(pair_key(M.enodes[2 * e], M.enodes[2 * e + 1]));
    if (it == count.end()) continue;
    prob->bnd[e] = bnd;
    n += it->second;  // the reference counts every matching line
  }
  if (nclassified) *nclassified = n;
  return 0;
}

int mom_problem_edge_bnd(const mom_problem *prob, int32_t dom, int32_t *bnd) {
  if (!prob || !bnd) return fail(MOM_ERR_ARG, "NULL argument");
  if (dom < -1 || dom >= (int32_t)prob->doms.size()) return fail(MOM_ERR_ARG, "domain index out of range");
  const HostMesh &m = dom < 0 ? prob->master : prob->doms[dom];
  for (int64_t e = 0; e < m.E(); e++) {
    const int64_t g = dom < 0 ? e : m.eparent[e];
    bnd[e] = prob->bnd.empty() ? kBndNone : prob->bnd[g];
  }
  return 0;
}

int mom_problem_destroy(mom_problem *prob) {
  delete prob;
  return 0;
}

int mom_problem_sizes(const mom_problem *prob, int32_t dom, int64_t *nedges_total, int64_t *nnodes, int64_t *nfaces,
                      int64_t *nedges) {
  if (!prob) return fail(MOM_ERR_ARG, "NULL argument");
  if (dom < -1 || dom >= (int32_t)prob->doms.size()) return fail(MOM_ERR_ARG, "domain index out of range");
  const HostMesh &m = dom < 0 ? prob->master : prob->doms[dom];
  if (nedges_total) *nedges_total = prob->master.E();
  if (nnodes) *nnodes = m.V();
  if (nfaces) *nfaces = m.F();
  if (nedges) *nedges = m.E();
  return 0;
}

int mom_problem_domain(const mom_problem *prob, int32_t dom, double *p, int32_t *fnodes, int32_t *fedges, double *fn,
                       double *fcp, double *fs, double *fm, double *farea, int32_t *efaces, int32_t *ebn,
                       double *elen, int32_t *eparent, double *avelen) {
  if (!prob) return fail(MOM_ERR_ARG, "NULL argument");
  if (dom < -1 || dom >= (int32_t)prob->doms.size()) return fail(MOM_ERR_ARG, "domain index out of range");
  const HostMesh &m = dom < 0 ? prob->master : prob->doms[dom];
  auto put = [](auto *dst, const auto &src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  put(p, m.p);
  put(fnodes, m.fnodes);
  put(fedges, m.fedges);
  put(fn, m.fn);
  put(fcp, m.fcp);
  put(fs, m.fs);
  put(fm, m.fm);
  put(farea, m.farea);
  put(efaces, m.efaces);
  put(ebn, m.ebn);
  put(elen, m.elen);
  put(eparent, m.eparent);
  if (avelen) *avelen = m.avelen;
  return 0;
}

int mom_problem_topology(const mom_problem *prob, int32_t dom, int32_t *enodes, int32_t *fid, int32_t *fparent,
                         int32_t *nparent) {
  if (!prob) return fail(MOM_ERR_ARG, "NULL argument");
  if (dom < -1 || dom >= (int32_t)prob->doms.size()) return fail(MOM_ERR_ARG, "domain index out of range");
  const HostMesh &m = dom < 0 ? prob->master : prob->doms[dom];
  auto put = [](int32_t *dst, const std::vector<int32_t> &src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  put(enodes, m.enodes);
  put(fid, m.fid);
  put(fparent, m.fparent);
  put(nparent, m.nparent);
  return 0;
}

int mom_problem_mesh_create(mom_ctx *ctx, const mom_problem *prob, int32_t dom, mom_mesh **mesh) {
  if (!ctx || !prob || !mesh) return fail(MOM_ERR_ARG, "NULL argument");
  if (dom < 0 || dom >= (int32_t)prob->doms.size()) return fail(MOM_ERR_ARG, "domain index out of range");
  const HostMesh &m = prob->doms[dom];
  return mom_mesh_create(ctx, m.V(), m.p.data(), m.F(), m.fnodes.data(), m.fedges.data(), m.fn.data(), m.fcp.data(),
                         m.fs.data(), m.fm.data(), m.farea.data(), m.E(), m.efaces.data(), m.ebn.data(),
                         m.elen.data(), m.eparent.data(), m.avelen, mesh);
}

}  // extern "C"
