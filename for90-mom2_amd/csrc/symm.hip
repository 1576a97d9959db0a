// Mirror-symmetry decomposition of the PMCHWT system.
//
// Reference algorithm (for90-mom2): the groups of `symm` (read_symm,
// interface.f90:160-210) built from group_id and group_mp (symmetry.f90:
// 45-104) by product_group (symmetry.f90:173-241); the representation
// matrices of sysmat_pmchwt (sysmat.f90:18-105) with the group-transformed
// source integrals (srcint.f90:155-224, 368-448) and near rule
// (sysmat.f90:233-260); the boundary conditions of edges on the mirror
// planes and the reduced systems (edge_bc, reduce_system, expand_solution,
// admissible_ga: bc.f90:28-257) solved per representation (solve_systems,
// solver.f90:327-357).
//
// Scope: the planes x = 0, y = 0, z = 0 (generators mxp, myp, mzp), the
// rotations about the z axis by multiples of 2 pi / n (generator rn,
// group_rz, symmetry.f90:105-156) and their products, groups of order up
// to 8.  The boundary classes of rotation cuts (rz1, rz2) and of periodic
// cells are rejected: the reference pairs edges (couple_index) only on the
// periodic boundaries (mesh.f90:195-258), so its rz1 condition
// (bc.f90:103-110) has no coupling to apply; rotation groups are for bodies
// that no cut crosses (n copies of a particle around the z axis), whose
// interior domains take the identity alone (admissible_ga).
//
// Each element g has an orthogonal Jacobian J_g.  The image fill of g
// integrates the test faces moved by J_g^T against the source faces
// (fill_domain with an Image): the reference integrates the test faces
// against the sources moved by J_g, and the Green's function and the dot
// products are invariant under the common rotation J_g^T; the K operator
// takes the factor det J_g.  The image blocks are added to every
// representation r's matrix with the field action ef[g][r] (complex for
// rotations; and det J_g where the reference puts it), each written
// straight into the reduced numbering of that representation: the rows and
// columns that edge_bc removes are never stored.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

Group identity_group() {
  Group g;
  g.n = 1;
  for (int i = 0; i < kMaxReps; i++) {
    for (int c = 0; c < 9; c++) g.J[i][c] = (c % 4 == 0) ? 1.0 : 0.0;
    g.detj[i] = 1.0;
    for (int r = 0; r < kMaxReps; r++) g.ef[i][r] = make_double2(1.0, 0.0);
    g.id[i] = kGidIdentity;
    g.genbits[i] = 1 << kGidIdentity;
  }
  return g;
}

bool jac_diagonal(const double *J) {
  return J[1] == 0.0 && J[2] == 0.0 && J[3] == 0.0 && J[5] == 0.0 && J[6] == 0.0 && J[7] == 0.0;
}

bool jac_involution(const double *J) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const double t = J[3 * i] * J[j] + J[3 * i + 1] * J[3 + j] + J[3 * i + 2] * J[6 + j];
      if (t != (i == j ? 1.0 : 0.0)) return false;
    }
  return true;
}

static double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// product_group (symmetry.f90:173-241): element k = n * |b| + m has the
// Jacobian J_n J_m (MATMUL), the union of the generator bits, the
// identifier of the non-identity factor when the other one is the identity
// (else 0), and the field action ef[k][k2] = ef_a[n][n2] ef_b[m][m2] with
// k2 = n2 |b| + m2.
static Group product(const Group &a, const Group &b) {
  Group g = identity_group();
  g.n = a.n * b.n;
  for (int n = 0; n < a.n; n++)
    for (int m = 0; m < b.n; m++) {
      const int k = n * b.n + m;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
#pragma clang fp contract(off)
          double t = a.J[n][3 * i] * b.J[m][j];
          t = t + a.J[n][3 * i + 1] * b.J[m][3 + j];
          t = t + a.J[n][3 * i + 2] * b.J[m][6 + j];
          g.J[k][3 * i + j] = t;
        }
      g.detj[k] = a.detj[n] * b.detj[m];
      g.genbits[k] = a.genbits[n] | b.genbits[m];
      int id = 0;
      if (a.id[n] == kGidIdentity && b.id[m] == kGidIdentity)
        id = kGidIdentity;
      else if (a.id[n] != 0 && a.id[n] != kGidIdentity && b.id[m] == kGidIdentity)
        id = a.id[n];
      else if (b.id[m] != 0 && b.id[m] != kGidIdentity && a.id[n] == kGidIdentity)
        id = b.id[m];
      g.id[k] = id;
      for (int n2 = 0; n2 < a.n; n2++)
        for (int m2 = 0; m2 < b.n; m2++) g.ef[k][n2 * b.n + m2] = cmul(a.ef[n][n2], b.ef[m][m2]);
    }
  return g;
}

// group_mp (symmetry.f90:61-104): {identity, mirror of plane np}, ef = [[1, 1], [1, -1]]
static Group mirror(int np) {
  Group mp = identity_group();
  mp.n = 2;
  mp.ef[1][1] = make_double2(-1.0, 0.0);
  mp.J[1][4 * (np - 1)] = -1.0;
  mp.detj[1] = -1.0;
  mp.id[1] = kGidIdentity + np;
  mp.genbits[1] = 1 << (kGidIdentity + np);
  return mp;
}

// group_rz (symmetry.f90:105-156): element k (0-based) is the counter-
// clockwise rotation about z by 2 pi k / n (matrix_rz, linalg.f90:324-335),
// det 1; element 0 is the identity and element n - 1 carries the rz
// identifier; every element but the identity has the rz generator bit; the
// field actions are ef[k][l] = conj(phase^(k l)), phase = exp(i 2 pi / n).
static Group rotation(int n) {
  Group g = identity_group();
  g.n = n;
  for (int k = 0; k < n; k++) {
    const double angle = 2.0 * MOM_PI * (double)k / (double)n;
    const double st = std::sin(angle), ct = std::cos(angle);
    const double J[9] = {ct, -st, 0.0, st, ct, 0.0, 0.0, 0.0, 1.0};
    for (int c = 0; c < 9; c++) g.J[k][c] = J[c];
    g.detj[k] = 1.0;
    g.id[k] = 0;
    g.genbits[k] = 1 << kGidRz;
  }
  g.id[0] = kGidIdentity;
  g.id[n - 1] = kGidRz;
  g.genbits[0] = 1 << kGidIdentity;
  const double angle = 2.0 * MOM_PI / (double)n;
  const double2 phase = make_double2(std::cos(angle), std::sin(angle));
  for (int k = 0; k < n; k++)
    for (int l = 0; l < n; l++) {
      double2 p = make_double2(1.0, 0.0);
      for (int e = 0; e < k * l; e++) p = cmul(p, phase);
      g.ef[k][l] = make_double2(p.x, -p.y);
    }
  return g;
}

int symm_group(int ngen, const int32_t *gens, Group *out) {
  if (ngen < 0 || ngen > 4 || (ngen > 0 && !gens)) return fail(MOM_ERR_ARG, "0 to 4 symmetry generators");
  Group g = identity_group();
  int seen = 0;
  for (int i = 0; i < ngen; i++) {
    const int code = gens[i];
    Group f;
    if (code >= 1 && code <= 3) {
      if (seen & (1 << code)) return fail(MOM_ERR_ARG, "mirror plane listed twice");
      seen |= 1 << code;
      f = mirror(code);
    } else if (code >= kGenRz + 2 && code <= kGenRz + kMaxReps) {
      if (seen & 1) return fail(MOM_ERR_ARG, "only one rotation generator");
      seen |= 1;
      f = rotation(code - kGenRz);
    } else {
      return fail(MOM_ERR_ARG, "generator must be 1 (mxp), 2 (myp), 3 (mzp) or 100 + n (rn, 2 <= n <= 8)");
    }
    if (g.n * f.n > kMaxReps) return fail(MOM_ERR_ARG, "the group has more than 8 elements");
    g = product(g, f);
  }
  // The decomposition uses one-dimensional representations, so the group
  // must be abelian: a rotation about z by other than pi does not commute
  // with the mirrors of x = 0 and y = 0 (with them it generates a dihedral
  // group), only with the mirror of z = 0.
  for (int i = 0; i < ngen; i++)
    if (gens[i] >= kGenRz + 3)
      for (int j = 0; j < ngen; j++)
        if (gens[j] == 1 || gens[j] == 2)
          return fail(MOM_ERR_ARG, "a rotation about z by 2 pi / n, n > 2, combines only with mzp (with mxp or myp "
                                   "the group is not abelian)");
  *out = g;
  return 0;
}

// edge_bc (bc.f90:28-118): an edge on the plane of a pure mirror element
// loses its J coefficient when Re(-ef) < 0.99 and its M coefficient when
// Re(-ef det J) < 0.99 (-1 marks a removed unknown).  Rotation elements
// act only on the rz1/rz2 boundary classes, which check_bnd rejects.
static void edge_bc(int64_t N, const int32_t *bnd, const Group &G, int rep, std::vector<int> &id) {
  id.assign(2 * N, 0);
  for (int64_t k = 0; k < N; k++)
    for (int n = 0; n < G.n; n++) {
      const int b = bnd[k];
      if ((G.id[n] == kGidMxp && b == kBndXplane) || (G.id[n] == kGidMyp && b == kBndYplane) ||
          (G.id[n] == kGidMzp && b == kBndZplane)) {
        const double gae = G.ef[n][rep].x, gah = gae * G.detj[n];
        if (-gae < 0.99) id[k] = -1;
        if (-gah < 0.99) id[k + N] = -1;
      }
    }
}

// Boundary classes other than none and the mirror planes (the periodic and
// rotation classes, mesh.f90:20-25) carry conditions this path does not
// build (bc.f90:55-110); a problem that uses them is rejected.
static int check_bnd(int64_t N, const int32_t *bnd) {
  for (int64_t k = 0; k < N; k++)
    if (bnd[k] < kBndNone || bnd[k] > kBndZplane)
      return fail(MOM_ERR_ARG, "edge " + std::to_string(k) + " has boundary class " + std::to_string(bnd[k]) +
                                   ": only none (1) and the mirror planes (2, 3, 4) are supported (rotation cuts "
                                   "rz1/rz2 and periodic boundaries are not)");
  return 0;
}

// reduce_system's numbering (bc.f90:170-193): the kept unknowns in order
static int64_t reduce_map(const std::vector<int> &id, std::vector<int> &map) {
  map.assign(id.size(), -1);
  int64_t dim = 0;
  for (size_t u = 0; u < id.size(); u++)
    if (id[u] != -1) map[u] = (int)dim++;
  return dim;
}

__global__ void set_diag_kernel(double2 *Z, int64_t ld, int64_t i0, int64_t n) {
  const int64_t i = i0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) Z[i + i * ld] = make_double2(1.0, 0.0);
}

int fill_system_symm(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *ri, int64_t N,
                     const int32_t *bnd, const Group &G, double omega, int32_t nq, const double *qw,
                     const double *qnodes, double2 *const *Z, const int64_t *ld, const int64_t *order) {
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  for (int d = 0; d < ndom; d++) {
    if (!doms[d]) return fail(MOM_ERR_ARG, "NULL domain mesh");
    if (doms[d]->ctx != ctx) return fail(MOM_ERR_ARG, "domain mesh belongs to another context");
    if (doms[d]->max_parent >= N) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  }
  MOM_TRY(check_bnd(N, bnd));
  std::vector<DevBuf> dmap(G.n);
  RepOut out;
  for (int r = 0; r < G.n; r++) {
    std::vector<int> id, map;
    edge_bc(N, bnd, G, r, id);
    const int64_t dim = reduce_map(id, map);
    // an odd order is padded by one decoupled unknown (unit diagonal)
    if (order[r] != dim + (dim & 1))
      return fail(MOM_ERR_ARG, "system order of representation " + std::to_string(r) + " must be " +
                                   std::to_string(dim + (dim & 1)));
    if ((rc = dmap[r].alloc(sizeof(int) * 2 * N))) return rc;
    MOM_TRY(copy_h2d(ctx, dmap[r].ptr, map.data(), sizeof(int) * 2 * N));
    MOM_HIP(hipMemset2DAsync(Z[r], ld[r] * sizeof(double2), 0, order[r] * sizeof(double2), order[r], ctx->stream));
    if (order[r] > dim)
      hipLaunchKernelGGL(set_diag_kernel, dim3(1), dim3(64), 0, ctx->stream, Z[r], ld[r], dim, order[r]);
    out.Z[r] = Z[r];
    out.ld[r] = ld[r];
    out.map[r] = dmap[r].as<int>();
  }
  for (int d = 0; d < ndom; d++) {
    // has_mesh_bnd of the domain (mesh.f90:86-100) through the edges' parents
    bool has[5] = {false, false, false, false, false};
    for (int64_t e = 0; e < doms[d]->E; e++) {
      const int b = bnd[doms[d]->h_eparent[e]];
      if (b >= 1 && b <= 4) has[b] = true;
    }
    const cd r(ri[2 * d], ri[2 * d + 1]);
    const cd zsi = (r * r) / (kEta0 * kEta0);
    for (int g = 0; g < G.n; g++) {
      // admissible_ga (bc.f90:237-257): the first (exterior) domain takes every
      // image, another one only the mirrors whose planes cut its surface
      // (and no rotation: its rz1 cut would be needed, which is rejected)
      if (d > 0 && (((G.genbits[g] >> kGidMxp) & 1 && !has[kBndXplane]) ||
                    ((G.genbits[g] >> kGidMyp) & 1 && !has[kBndYplane]) ||
                    ((G.genbits[g] >> kGidMzp) & 1 && !has[kBndZplane]) || ((G.genbits[g] >> kGidRz) & 1)))
        continue;
      FillTarget t;
      t.Noff = N;
      t.mode = 0;
      t.zsi = make_double2(zsi.real(), zsi.imag());
      t.nrep = G.n;
      t.detj = G.detj[g];
      t.rep = out;
      for (int k = 0; k < G.n; k++) t.rep.e[k] = G.ef[g][k];
      Image img;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) img.M[3 * i + j] = G.J[g][3 * j + i];
      img.detj = G.detj[g];
      img.diag = jac_diagonal(G.J[g]);
      img.invol = jac_involution(G.J[g]);
      const bool ident = img.diag && G.J[g][0] == 1.0 && G.J[g][4] == 1.0 && G.J[g][8] == 1.0;
      if ((rc = fill_domain(ctx, doms[d], rule, omega, r, t, ident ? nullptr : &img))) return rc;
    }
  }
  // the maps must outlive the queued kernels
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace mom

using namespace mom;

extern "C" {

int mom_symm_group(int32_t ngen, const int32_t *planes, int32_t *nga, double *ef, double *detj, double *jdiag,
                   int32_t *id) {
  if (!nga) return fail(MOM_ERR_ARG, "NULL argument");
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  for (int g = 0; g < G.n; g++)
    if (!jac_diagonal(G.J[g]))
      return fail(MOM_ERR_ARG, "a rotation group has complex field actions and full Jacobians: use mom_symm_group_c");
  *nga = G.n;
  for (int g = 0; g < G.n; g++) {
    if (ef)
      for (int r = 0; r < G.n; r++) ef[g * G.n + r] = G.ef[g][r].x;
    if (detj) detj[g] = G.detj[g];
    if (jdiag)
      for (int c = 0; c < 3; c++) jdiag[3 * g + c] = G.J[g][4 * c];
    if (id) id[g] = G.id[g];
  }
  return 0;
}

int mom_symm_group_c(int32_t ngen, const int32_t *gens, int32_t *nga, double *ef, double *detj, double *jmat,
                     int32_t *id) {
  if (!nga) return fail(MOM_ERR_ARG, "NULL argument");
  Group G;
  MOM_TRY(symm_group(ngen, gens, &G));
  *nga = G.n;
  for (int g = 0; g < G.n; g++) {
    if (ef)
      for (int r = 0; r < G.n; r++) {
        ef[2 * (g * G.n + r)] = G.ef[g][r].x;
        ef[2 * (g * G.n + r) + 1] = G.ef[g][r].y;
      }
    if (detj) detj[g] = G.detj[g];
    if (jmat)
      for (int c = 0; c < 9; c++) jmat[9 * g + c] = G.J[g][c];
    if (id) id[g] = G.id[g];
  }
  return 0;
}

int mom_symm_edge_bc(int64_t nedges_total, const int32_t *bnd, int32_t ngen, const int32_t *planes, int32_t rep,
                     int32_t *id, int64_t *dim) {
  if (!bnd || !id || !dim) return fail(MOM_ERR_ARG, "NULL argument");
  if (nedges_total < 1) return fail(MOM_ERR_ARG, "nedges_total must be >= 1");
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  if (rep < 0 || rep >= G.n) return fail(MOM_ERR_ARG, "representation index out of range");
  MOM_TRY(check_bnd(nedges_total, bnd));
  std::vector<int> v, map;
  edge_bc(nedges_total, bnd, G, rep, v);
  *dim = reduce_map(v, map);
  std::copy(v.begin(), v.end(), id);
  return 0;
}

}  // extern "C"
