// Adjoint sensitivities of the PMCHWT solution (no counterpart in the
// reference; its forward quantities are sysmat_pmchwt, sysmat.f90:18-105,
// and cs_prtsrf, cs.f90:14-104).
//
// For an observable sigma(x) of the solution of Z x = b, with
// dsigma = Re(g^T dx), the derivative with respect to a parameter p of Z is
//   dsigma/dp = -Re(lambda^T (dZ/dp) x),  Z^T lambda = g,
// and lambda = conj(mu) with Z^H mu = conj(g) (mom_system_solve_adjoint).
// This file provides the two pieces around that solve:
//  - mom_cs_abs_gradient: sigma_abs of cs_prtsrf and its gradient g;
//  - mom_pmchwt_deps_apply: (dZ/deps_d) x for the relative permittivity
//    eps_d = ri_d^2 of one domain, from a fill of the k-derivatives of that
//    domain's operators (dk/deps = k / (2 eps)) plus the explicit eps
//    dependence of the H prefactor (1/eps) and of Zsi = eps/eta0^2.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

// sigma_abs = (c/2) Re(xM^T W conj(xJ)) with, per face,
// W_ab = A sum_r w_r ((n x f_a) x (n x f_b)) . n  (cs.f90:66-94 with the
// total fields Et = sum xM (n x f), Ht = -sum xJ (n x f)).  Per face side:
// gM[3f+a] = (c/2) sum_b W_ab conj(xJ_b), gJ[3f+b] = (c/2) sum_a W_ab conj(xM_a),
// part[f] = (c/2) Re(sum_ab xM_a W_ab conj(xJ_b)).
struct CsGradArgs {
  int F, Q;
  double w[kMaxQ];
  const double *qp, *fn, *area;
  const int *g;
  const double *coef, *v;
  const double2 *x;
  int64_t Nt;
  double half_c;
  double2 *gM, *gJ;
  double *part;
};

__global__ __launch_bounds__(128) void cs_abs_grad_kernel(CsGradArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const double n[3] = {a.fn[3 * f], a.fn[3 * f + 1], a.fn[3 * f + 2]};
  double W[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int t = 0; t < a.Q; t++) {
    const double *r = a.qp + ((int64_t)f * a.Q + t) * 3;
    double u[3][3];
    for (int q = 0; q < 3; q++) {
      const int s = 3 * f + q;
      double fv[3];
      for (int c = 0; c < 3; c++) fv[c] = a.coef[s] * (r[c] - a.v[3 * s + c]);
      u[q][0] = n[1] * fv[2] - n[2] * fv[1];
      u[q][1] = n[2] * fv[0] - n[0] * fv[2];
      u[q][2] = n[0] * fv[1] - n[1] * fv[0];
    }
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) {
        const double cx = u[p][1] * u[q][2] - u[p][2] * u[q][1];
        const double cy = u[p][2] * u[q][0] - u[p][0] * u[q][2];
        const double cz = u[p][0] * u[q][1] - u[p][1] * u[q][0];
        W[p][q] += a.w[t] * (cx * n[0] + cy * n[1] + cz * n[2]);
      }
  }
  const double A = a.area[f];
  double2 xm[3], xj[3];
  for (int q = 0; q < 3; q++) {
    xj[q] = a.x[a.g[3 * f + q]];
    xm[q] = a.x[a.Nt + a.g[3 * f + q]];
  }
  double sre = 0;
  for (int p = 0; p < 3; p++) {
    double gmr = 0, gmi = 0, gjr = 0, gji = 0;
    for (int q = 0; q < 3; q++) {
      const double wpq = W[p][q] * A, wqp = W[q][p] * A;
      // conj(xJ_q) W_pq and conj(xM_q) W_qp
      gmr += wpq * xj[q].x;
      gmi -= wpq * xj[q].y;
      gjr += wqp * xm[q].x;
      gji -= wqp * xm[q].y;
    }
    a.gM[3 * f + p] = make_double2(a.half_c * gmr, a.half_c * gmi);
    a.gJ[3 * f + p] = make_double2(a.half_c * gjr, a.half_c * gji);
    sre += xm[p].x * gmr - xm[p].y * gmi;
  }
  a.part[f] = a.half_c * sre;
}

// y = M x for an nr x nc complex128 matrix (ld): 32 column chunks per row
// block (partial sums in a fixed order, then one reduction), so that a
// block of a few thousand rows still spreads over the whole device
constexpr int kMvChunks = 32;

__global__ __launch_bounds__(256) void matvec_part_kernel(const double2 *__restrict__ M, int64_t ld, int64_t nr,
                                                          int64_t nc, const double2 *__restrict__ x, double2 *part) {
  const int64_t cw = (nc + kMvChunks - 1) / kMvChunks;
  const int64_t c0 = blockIdx.y * cw, c1 = std::min<int64_t>(nc, c0 + cw);
  const int64_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nr) return;
  double ar = 0, ai = 0;
  for (int64_t c = c0; c < c1; c++) {
    const double2 m = M[r + c * ld], v = x[c];
    ar += m.x * v.x - m.y * v.y;
    ai += m.x * v.y + m.y * v.x;
  }
  part[blockIdx.y * nr + r] = make_double2(ar, ai);
}

__global__ void matvec_reduce_kernel(const double2 *__restrict__ part, int64_t nr, double2 *y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nr) return;
  double ar = 0, ai = 0;
  for (int k = 0; k < kMvChunks; k++) {
    ar += part[k * nr + r].x;
    ai += part[k * nr + r].y;
  }
  y[r] = make_double2(ar, ai);
}

int matvec(mom_ctx *ctx, const double2 *M, int64_t ld, int64_t nr, int64_t nc, const double2 *x, double2 *y) {
  DevBuf part;
  MOM_TRY(part.alloc(sizeof(double2) * nr * kMvChunks));
  hipLaunchKernelGGL(matvec_part_kernel, dim3((unsigned)((nr + 255) / 256), kMvChunks), dim3(256), 0, ctx->stream, M,
                     ld, nr, nc, x, part.as<double2>());
  hipLaunchKernelGGL(matvec_reduce_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, ctx->stream,
                     part.as<double2>(), nr, y);
  MOM_HIP(hipGetLastError());
  // the partial sums must outlive the kernels
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// y = M^T x (no conjugation): one wave per column of the column-major M,
// lanes striding down the column, the wave's partial sums added in one
// fixed order (deterministic).
__global__ __launch_bounds__(256) void matvec_t_kernel(const double2 *__restrict__ M, int64_t ld, int64_t nr,
                                                       int64_t nc, const double2 *__restrict__ x, double2 *y) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= nc) return;
  const double2 *col = M + c * ld;
  double ar = 0, ai = 0;
  for (int64_t r = lane; r < nr; r += 64) {
    const double2 m = col[r], v = x[r];
    ar += m.x * v.x - m.y * v.y;
    ai += m.x * v.y + m.y * v.x;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o);
    ai += __shfl_xor(ai, o);
  }
  if (lane == 0) y[c] = make_double2(ar, ai);
}

}  // namespace mom

using namespace mom;

// y = Z_d x or Z_d^T x for one domain's blocks of the PMCHWT matrix
// (sysmat.f90:56-98 for that domain alone) in global numbering: the block
// is filled in the domain's local numbering, applied to x at the domain's
// unknowns, and the product returned at them (zero elsewhere).
extern "C" int mom_domain_block_apply(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega,
                                      double ri_re, double ri_im, int32_t trans, int32_t nq, const double *qw,
                                      const double *qnodes, const double *x, double *y) {
  if (!ctx || !mesh || !x || !y) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total)
    return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  if (trans != 0 && trans != 1) return fail(MOM_ERR_ARG, "trans must be 0 or 1");
  const cd ri(ri_re, ri_im);
  if (ri == cd(0.0)) return fail(MOM_ERR_ARG, "ri must be non-zero");
  (void)hipSetDevice(ctx->device);
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  const int64_t Nt = nedges_total, E = mesh->E;
  const cd zsi = (ri * ri) / (kEta0 * kEta0);
  DevBuf Z, dx, dy;
  int rc;
  if ((rc = Z.alloc(sizeof(double2) * 4 * E * E)) || (rc = dx.alloc(sizeof(double2) * 2 * E)) ||
      (rc = dy.alloc(sizeof(double2) * 2 * E)))
    return rc;
  MOM_HIP(hipMemsetAsync(Z.ptr, 0, Z.bytes, ctx->stream));
  FillTarget t;
  t.Z = Z.as<double2>();
  t.ld = 2 * E;
  t.Noff = E;
  t.mode = 0;
  t.local = 1;
  t.zsi = make_double2(zsi.real(), zsi.imag());
  MOM_TRY(fill_domain(ctx, mesh, rule, omega, ri, t));
  std::vector<cd> xl(2 * E), yl(2 * E);
  for (int64_t i = 0; i < E; i++) {
    const int64_t gi = mesh->h_eparent[i];
    xl[i] = cd(x[2 * gi], x[2 * gi + 1]);
    xl[E + i] = cd(x[2 * (Nt + gi)], x[2 * (Nt + gi) + 1]);
  }
  MOM_TRY(copy_h2d(ctx, dx.ptr, xl.data(), sizeof(double2) * 2 * E));
  if (trans) {
    hipLaunchKernelGGL(matvec_t_kernel, dim3((unsigned)((2 * E + 3) / 4)), dim3(256), 0, ctx->stream, Z.as<double2>(),
                       2 * E, 2 * E, 2 * E, dx.as<double2>(), dy.as<double2>());
    MOM_HIP(hipGetLastError());
  } else {
    MOM_TRY(matvec(ctx, Z.as<double2>(), 2 * E, 2 * E, 2 * E, dx.as<double2>(), dy.as<double2>()));
  }
  MOM_TRY(copy_d2h(ctx, yl.data(), dy.ptr, sizeof(double2) * 2 * E));
  for (int64_t i = 0; i < 4 * Nt; i++) y[i] = 0.0;
  for (int64_t i = 0; i < E; i++) {
    const int64_t gi = mesh->h_eparent[i];
    y[2 * gi] = yl[i].real();
    y[2 * gi + 1] = yl[i].imag();
    y[2 * (Nt + gi)] = yl[E + i].real();
    y[2 * (Nt + gi) + 1] = yl[E + i].imag();
  }
  return 0;
}

extern "C" {

int mom_cs_abs_gradient(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                        double ri_im, const double *x, int32_t nq, const double *qw, const double *qnodes, double *g,
                        double *cabs) {
  (void)omega;
  (void)ri_im;
  if (!ctx || !mesh || !x || !g) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  if (!(ri_re != 0)) return fail(MOM_ERR_ARG, "Re(ri) must be non-zero");
  (void)hipSetDevice(ctx->device);
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  MOM_TRY(compute_quad_points(ctx, mesh, rule));
  const int64_t Nt = nedges_total, F = mesh->F;
  DevBuf dx, dgm, dgj, dpart;
  int rc;
  if ((rc = dx.alloc(sizeof(double2) * 2 * Nt)) || (rc = dgm.alloc(sizeof(double2) * 3 * F)) ||
      (rc = dgj.alloc(sizeof(double2) * 3 * F)) || (rc = dpart.alloc(sizeof(double) * F)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dx.ptr, x, sizeof(double2) * 2 * Nt));
  CsGradArgs a;
  a.F = (int)F;
  a.Q = rule.Q;
  for (int i = 0; i < kMaxQ; i++) a.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  a.qp = ctx->qp.as<double>();
  a.fn = mesh->fn.as<double>();
  a.area = mesh->farea.as<double>();
  a.g = mesh->s_g.as<int>();
  a.coef = mesh->s_coef.as<double>();
  a.v = mesh->s_v.as<double>();
  a.x = dx.as<double2>();
  a.Nt = Nt;
  // cs.f90:38: c = 2 / (eps0 c0 Re(ri))
  a.half_c = 1.0 / (kEps0 * kC0 * ri_re);
  a.gM = dgm.as<double2>();
  a.gJ = dgj.as<double2>();
  a.part = dpart.as<double>();
  hipLaunchKernelGGL(cs_abs_grad_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, ctx->stream, a);
  MOM_HIP(hipGetLastError());
  std::vector<double2> gm(3 * F), gj(3 * F);
  std::vector<double> part(F);
  MOM_TRY(copy_d2h(ctx, gm.data(), dgm.ptr, sizeof(double2) * 3 * F));
  MOM_TRY(copy_d2h(ctx, gj.data(), dgj.ptr, sizeof(double2) * 3 * F));
  MOM_TRY(copy_d2h(ctx, part.data(), dpart.ptr, sizeof(double) * F));
  for (int64_t i = 0; i < 4 * Nt; i++) g[i] = 0.0;
  // per face side into the global unknowns, faces in order
  for (int64_t s = 0; s < 3 * F; s++) {
    const int64_t e = mesh->h_eparent[mesh->h_fedges[s]];
    g[2 * e] += gj[s].x;
    g[2 * e + 1] += gj[s].y;
    g[2 * (Nt + e)] += gm[s].x;
    g[2 * (Nt + e) + 1] += gm[s].y;
  }
  if (cabs) {
    double acc = 0;
    for (int64_t f = 0; f < F; f++) acc += part[f];
    *cabs = acc;
  }
  return 0;
}

int mom_pmchwt_deps_apply(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                          double ri_im, int32_t nq, const double *qw, const double *qnodes, const double *x,
                          double *y) {
  if (!ctx || !mesh || !x || !y) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  const cd ri(ri_re, ri_im);
  if (ri == cd(0.0)) return fail(MOM_ERR_ARG, "ri must be non-zero");
  (void)hipSetDevice(ctx->device);
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  const int64_t Nt = nedges_total, E = mesh->E;
  const cd eps = ri * ri, k = ri * omega / kC0;
  const cd zsi = eps / (kEta0 * kEta0);
  DevBuf P, D, H, dxl, dyp, dyd, dyh, dyhm;
  int rc;
  if ((rc = P.alloc(sizeof(double2) * 4 * E * E)) || (rc = D.alloc(sizeof(double2) * E * E)) ||
      (rc = H.alloc(sizeof(double2) * E * E)) || (rc = dxl.alloc(sizeof(double2) * 2 * E)) ||
      (rc = dyp.alloc(sizeof(double2) * 2 * E)) || (rc = dyd.alloc(sizeof(double2) * E)) ||
      (rc = dyh.alloc(sizeof(double2) * E)) || (rc = dyhm.alloc(sizeof(double2) * E)))
    return rc;
  MOM_HIP(hipMemsetAsync(P.ptr, 0, P.bytes, ctx->stream));
  MOM_HIP(hipMemsetAsync(D.ptr, 0, D.bytes, ctx->stream));
  MOM_HIP(hipMemsetAsync(H.ptr, 0, H.bytes, ctx->stream));
  FillTarget t;
  t.Z = P.as<double2>();
  t.ld = 2 * E;
  t.Noff = E;
  t.mode = 0;
  t.local = 1;
  t.dk = 1;
  t.zsi = make_double2(zsi.real(), zsi.imag());
  MOM_TRY(fill_domain(ctx, mesh, rule, omega, ri, t));
  FillTarget td;
  td.Z = D.as<double2>();
  td.ld = E;
  td.mode = 1;
  MOM_TRY(fill_domain(ctx, mesh, rule, omega, ri, td));
  td.Z = H.as<double2>();
  td.mode = 2;
  MOM_TRY(fill_domain(ctx, mesh, rule, omega, ri, td));
  // local x = (xJ, xM) at the domain's unknowns
  std::vector<cd> xl(2 * E);
  for (int64_t i = 0; i < E; i++) {
    const int64_t gi = mesh->h_eparent[i];
    xl[i] = cd(x[2 * gi], x[2 * gi + 1]);
    xl[E + i] = cd(x[2 * (Nt + gi)], x[2 * (Nt + gi) + 1]);
  }
  MOM_TRY(copy_h2d(ctx, dxl.ptr, xl.data(), sizeof(double2) * 2 * E));
  const double2 *xJ = dxl.as<double2>(), *xM = dxl.as<double2>() + E;
  MOM_TRY(matvec(ctx, P.as<double2>(), 2 * E, 2 * E, 2 * E, xJ, dyp.as<double2>()));
  MOM_TRY(matvec(ctx, D.as<double2>(), E, E, E, xM, dyd.as<double2>()));
  MOM_TRY(matvec(ctx, H.as<double2>(), E, E, E, xJ, dyh.as<double2>()));
  MOM_TRY(matvec(ctx, H.as<double2>(), E, E, E, xM, dyhm.as<double2>()));
  std::vector<cd> yp(2 * E), ydm(E), yhj(E), yhm(E);
  MOM_TRY(copy_d2h(ctx, yp.data(), dyp.ptr, sizeof(double2) * 2 * E));
  MOM_TRY(copy_d2h(ctx, ydm.data(), dyd.ptr, sizeof(double2) * E));
  MOM_TRY(copy_d2h(ctx, yhj.data(), dyh.ptr, sizeof(double2) * E));
  MOM_TRY(copy_d2h(ctx, yhm.data(), dyhm.ptr, sizeof(double2) * E));
  // dZ/deps x = (k/(2 eps)) P x + [-H xJ / eps ; (D + H) xM / eta0^2 - zsi H xM / eps]
  const cd dk = k / (2.0 * eps);
  for (int64_t i = 0; i < 4 * Nt; i++) y[i] = 0.0;
  for (int64_t i = 0; i < E; i++) {
    const int64_t gi = mesh->h_eparent[i];
    const cd a = dk * yp[i] - yhj[i] / eps;
    const cd b = dk * yp[E + i] + (ydm[i] + yhm[i]) / (kEta0 * kEta0) - zsi * yhm[i] / eps;
    y[2 * gi] += a.real();
    y[2 * gi + 1] += a.imag();
    y[2 * (Nt + gi)] += b.real();
    y[2 * (Nt + gi) + 1] += b.imag();
  }
  return 0;
}


// The vertex-coordinate terms of an adjoint sensitivity: dvert[3V] (host,
// this mesh's nodes) += Re d/dP (lambda^T Z_mesh x) for the mesh's blocks of
// the PMCHWT matrix (mom_pmchwt_vertex_adjoint), or += d/dP (sigma_abs at
// fixed x + Re lambda^T q) for the exterior mesh and one plane wave
// (mom_cs_abs_vertex_explicit).
static int vertex_common(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, const double *x, const double *lam,
                         double *dvert, int32_t nq, const double *qw, const double *qnodes, QuadRule &rule, DevBuf &dx,
                         DevBuf &dl, DevBuf &dv) {
  if (!ctx || !mesh || !x || !lam || !dvert) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total)
    return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  (void)hipSetDevice(ctx->device);
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  int rc;
  if ((rc = dx.alloc(sizeof(double2) * 2 * nedges_total)) || (rc = dl.alloc(sizeof(double2) * 2 * nedges_total)) ||
      (rc = dv.alloc(sizeof(double) * 9 * mesh->F)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dx.ptr, x, sizeof(double2) * 2 * nedges_total));
  MOM_TRY(copy_h2d(ctx, dl.ptr, lam, sizeof(double2) * 2 * nedges_total));
  return 0;
}

// The faces' values (dv: [F][9], every face's own local nodes) summed into
// the nodes in face order, so that the node gradient does not depend on
// the order in which the device finished.
static int vertex_finish(mom_ctx *ctx, const mom_mesh *mesh, DevBuf &dv, double *dvert) {
  std::vector<double> h(9 * mesh->F);
  MOM_TRY(copy_d2h(ctx, h.data(), dv.ptr, sizeof(double) * h.size()));
  for (int64_t f = 0; f < mesh->F; f++)
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) dvert[3 * mesh->h_fnodes[3 * f + v] + c] += h[9 * f + 3 * v + c];
  return 0;
}

// test hook: far pairs by differences as well (mom_test_vertex_fd)
static int g_vertex_fd = 0;

int mom_test_vertex_fd(int32_t enable) {
  g_vertex_fd = enable ? 1 : 0;
  return 0;
}

int mom_pmchwt_vertex_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                              double ri_im, int32_t nq, const double *qw, const double *qnodes, const double *x,
                              const double *lam, double *dvert) {
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  QuadRule rule;
  DevBuf dx, dl, dv;
  MOM_TRY(vertex_common(ctx, mesh, nedges_total, x, lam, dvert, nq, qw, qnodes, rule, dx, dl, dv));
  MOM_TRY(vertex_adjoint_fill(ctx, mesh, rule, omega, cd(ri_re, ri_im), nedges_total, dx.as<double2>(),
                              dl.as<double2>(), dv.as<double>(), g_vertex_fd));
  return vertex_finish(ctx, mesh, dv, dvert);
}

int mom_cs_abs_vertex_explicit(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                               double ri_im, const double *x, const double *lam, double theta, double phi, double psi,
                               int32_t nq, const double *qw, const double *qnodes, double *dvert) {
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  if (!(ri_re > 0)) return fail(MOM_ERR_ARG, "Re(ri) must be > 0");
  QuadRule rule;
  DevBuf dx, dl, dv;
  MOM_TRY(vertex_common(ctx, mesh, nedges_total, x, lam, dvert, nq, qw, qnodes, rule, dx, dl, dv));
  MOM_TRY(vertex_adjoint_faces(ctx, mesh, rule, omega, cd(ri_re, ri_im), nedges_total, dx.as<double2>(),
                               dl.as<double2>(), theta, phi, psi, dv.as<double>()));
  return vertex_finish(ctx, mesh, dv, dvert);
}

}  // extern "C"
