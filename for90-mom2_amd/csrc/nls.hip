// Second-harmonic generation from a surface chi2 (centrosymmetric media),
// identity group, free space.
//
// Reference: nlsurf_coef (nlsurf.f90:137-338) with efield, Pnls_invmap and
// Pnls_frag (21-126), rwg_moments (rwgf.f90:148-221) and rwg_exp (278-298);
// sysmat_pmchwt_nls (sysmat.f90:112-229); the caller is the centrosymmetric
// branch of solve_batch (solver.f90:143-231).
//
// nlsurf_coef of one domain mesh:
//  1. per face and quadrature point, the surface field of the fundamental
//     solution, E = n x M + n divJ/(i omega eps), and the polarisation
//     P = eps0 (chi_nnn En^2 + chi_ntt Et.Et) n + 2 eps0 chi_ttn Et En;
//     projected on the face's RWG functions: coef = <div f, -Pn> A/epsp,
//     srcJ = <f, Pt> A, srcH = <f, Pt x n> A, summed per edge over its faces
//     (one thread per face, then one per edge);
//  2. Gram solves F [coef srcJ] = rhs with F = <f_m, f_n>.  F is real,
//     symmetric positive definite and has at most five entries per row; the
//     reference factors it densely (ZGETRF).  Here it is kept sparse (CSR
//     built on the host, values by one thread per entry) and solved by
//     Jacobi-preconditioned conjugate gradients on real and imaginary parts,
//     one workgroup per real right-hand side;
//  3. srcM = <f, n x sum coef f> A per edge, and a second Gram solve.
// sysmat_pmchwt_nls: the SH matrix is the PMCHWT matrix at the SH frequency
// with the SH indices; the source loses Z_d [srcJ; srcM] for each nonlinear
// domain d, with Z_d that domain's own PMCHWT block.  That block is filled
// once in local numbering, multiplied, and added into the global matrix.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

constexpr int kCgThreads = 1024;
constexpr double kCgTol = 1e-15;  // recursive residual, relative to the right-hand side (2-norm)

struct NlsFaceArgs {
  int F, Q;
  double w[kMaxQ];
  const double *qp;       // [F][Q][3]
  const double *fn;       // [F][3]
  const double *area;     // [F]
  const int *g;           // global unknown of each face side [3F]
  const double *coef;     // sign L/(2A) [3F]
  const double *div;      // sign L/A [3F]
  const double *v;        // free vertex [3F][3]
  const double2 *x;       // fundamental solution [2 Nt] (J then M)
  int64_t Nt;
  double2 en_fac;         // 1 / (i omega eps), eps = eps0 ri^2
  double2 chi[3];         // nnn, ntt, ttn
  const double2 *epsp;    // [F]
  double2 *side_out;      // [3F][3]: int1 A/epsp, int3 A, int2 A
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// nlsurf.f90:21-126 and 186-217: one thread per face
__global__ __launch_bounds__(128) void nls_stage1_kernel(NlsFaceArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const double n[3] = {a.fn[3 * f], a.fn[3 * f + 1], a.fn[3 * f + 2]};
  double2 xj[3], xm[3];
  double cf[3], dv[3], vv[3][3];
  for (int q = 0; q < 3; q++) {
    const int s = 3 * f + q;
    xj[q] = a.x[a.g[s]];
    xm[q] = a.x[a.Nt + a.g[s]];
    cf[q] = a.coef[s];
    dv[q] = a.div[s];
    for (int c = 0; c < 3; c++) vv[q][c] = a.v[3 * s + c];
  }
  // div J and en are the same at every point of the face
  double2 divj = make_double2(0, 0);
  for (int q = 0; q < 3; q++) divj = cadd(divj, cscale(xj[q], dv[q]));
  const double2 en = cmul(divj, a.en_fac);
  double2 i1[3], i2[3], i3[3];
  for (int q = 0; q < 3; q++) i1[q] = i2[q] = i3[q] = make_double2(0, 0);
  for (int t = 0; t < a.Q; t++) {
    const double *r = a.qp + ((int64_t)f * a.Q + t) * 3;
    double fv[3][3];
    for (int q = 0; q < 3; q++)
      for (int c = 0; c < 3; c++) fv[q][c] = cf[q] * (r[c] - vv[q][c]);
    double2 ms[3];
    for (int c = 0; c < 3; c++) {
      ms[c] = make_double2(0, 0);
      for (int q = 0; q < 3; q++) ms[c] = cadd(ms[c], cscale(xm[q], fv[q][c]));
    }
    // e = n x ms + n en
    double2 e[3];
    e[0] = cadd(csub(cscale(ms[2], n[1]), cscale(ms[1], n[2])), cscale(en, n[0]));
    e[1] = cadd(csub(cscale(ms[0], n[2]), cscale(ms[2], n[0])), cscale(en, n[1]));
    e[2] = cadd(csub(cscale(ms[1], n[0]), cscale(ms[0], n[1])), cscale(en, n[2]));
    // polarisation (nlsurf.f90:91-100), dotc without conjugation
    const double2 enn = cadd(cadd(cscale(e[0], n[0]), cscale(e[1], n[1])), cscale(e[2], n[2]));
    double2 et[3];
    for (int c = 0; c < 3; c++) et[c] = csub(e[c], cscale(enn, n[c]));
    const double2 ett = cadd(cadd(cmul(et[0], et[0]), cmul(et[1], et[1])), cmul(et[2], et[2]));
    const double2 pn0 = cscale(cadd(cmul(a.chi[0], cmul(enn, enn)), cmul(a.chi[1], ett)), kEps0);
    const double2 tfac = cmul(cscale(a.chi[2], 2.0 * kEps0), enn);
    double2 P[3];
    for (int c = 0; c < 3; c++) P[c] = cadd(cmul(et[c], tfac), cscale(pn0, n[c]));
    // normal and tangential parts (nlsurf.f90:189-191)
    const double2 pn = cadd(cadd(cscale(P[0], n[0]), cscale(P[1], n[1])), cscale(P[2], n[2]));
    double2 pt[3];
    for (int c = 0; c < 3; c++) pt[c] = csub(P[c], cscale(pn, n[c]));
    // pt x n
    double2 ptn[3];
    ptn[0] = csub(cscale(pt[1], n[2]), cscale(pt[2], n[1]));
    ptn[1] = csub(cscale(pt[2], n[0]), cscale(pt[0], n[2]));
    ptn[2] = csub(cscale(pt[0], n[1]), cscale(pt[1], n[0]));
    const double w = a.w[t];
    for (int q = 0; q < 3; q++) {
      i1[q] = csub(i1[q], cscale(pn, w * dv[q]));
      double2 d2 = make_double2(0, 0), d3 = make_double2(0, 0);
      for (int c = 0; c < 3; c++) {
        d2 = cadd(d2, cscale(ptn[c], fv[q][c]));
        d3 = cadd(d3, cscale(pt[c], fv[q][c]));
      }
      i2[q] = cadd(i2[q], cscale(d2, w));
      i3[q] = cadd(i3[q], cscale(d3, w));
    }
  }
  const double A = a.area[f];
  const double2 ep = a.epsp[f];
  for (int q = 0; q < 3; q++) {
    double2 *o = a.side_out + 3 * (3 * f + q);
    o[0] = cdiv(cscale(i1[q], A), ep);
    o[1] = cscale(i3[q], A);
    o[2] = cscale(i2[q], A);
  }
}

// Per edge: the sums over its two face sides, scaled as nlsurf.f90:209-213.
// B[4][E] (real): Re coef, Im coef, Re srcJ, Im srcJ (the first Gram
// solve's right-hand sides); coef0 and sh keep coef and the scaled H source.
__global__ void nls_gather1_kernel(int E, const int *__restrict__ eside, const double2 *__restrict__ side_out,
                                   double omegash, double *B, double2 *coef0, double2 *srch) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double2 c = make_double2(0, 0), j = make_double2(0, 0), h = make_double2(0, 0);
  for (int k = 0; k < 2; k++) {
    const int s = eside[2 * e + k];
    if (s < 0) continue;
    c = cadd(c, side_out[3 * s]);
    j = cadd(j, side_out[3 * s + 1]);
    h = cadd(h, side_out[3 * s + 2]);
  }
  // srcJ = -srcJ i omegash; srcH = 0.5 i omegash srcH
  const double2 jj = make_double2(j.y * omegash, -j.x * omegash);
  B[e] = c.x;
  B[E + e] = c.y;
  B[2 * E + e] = jj.x;
  B[3 * E + e] = jj.y;
  coef0[e] = c;
  srch[e] = make_double2(-0.5 * omegash * h.y, 0.5 * omegash * h.x);
}

// nlsurf.f90:278-299: per face side A <f_q, n x sum_t coef_t f_t>, with coef
// the first solve's solution (X rows 0, 1 = real, imaginary parts)
struct NlsStage2Args {
  int F, Q, E;
  double w[kMaxQ];
  const double *qp, *fn, *area;
  const int *l;  // local edge of each face side
  const double *coef, *v;
  const double *X;  // [>=2][E]
  double2 *side_out;  // [3F]
};

__global__ __launch_bounds__(128) void nls_stage2_kernel(NlsStage2Args a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const double n[3] = {a.fn[3 * f], a.fn[3 * f + 1], a.fn[3 * f + 2]};
  double cf[3], vv[3][3];
  double2 cc[3];
  for (int q = 0; q < 3; q++) {
    const int s = 3 * f + q;
    const int le = a.l[s];
    cc[q] = make_double2(a.X[le], a.X[a.E + le]);
    cf[q] = a.coef[s];
    for (int c = 0; c < 3; c++) vv[q][c] = a.v[3 * s + c];
  }
  double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int t = 0; t < a.Q; t++) {
    const double *r = a.qp + ((int64_t)f * a.Q + t) * 3;
    double fv[3][3];
    for (int q = 0; q < 3; q++)
      for (int c = 0; c < 3; c++) fv[q][c] = cf[q] * (r[c] - vv[q][c]);
    double2 g[3];
    for (int c = 0; c < 3; c++) {
      g[c] = make_double2(0, 0);
      for (int q = 0; q < 3; q++) g[c] = cadd(g[c], cscale(cc[q], fv[q][c]));
    }
    double2 h[3];
    h[0] = csub(cscale(g[2], n[1]), cscale(g[1], n[2]));
    h[1] = csub(cscale(g[0], n[2]), cscale(g[2], n[0]));
    h[2] = csub(cscale(g[1], n[0]), cscale(g[0], n[1]));
    for (int q = 0; q < 3; q++) {
      double2 d = make_double2(0, 0);
      for (int c = 0; c < 3; c++) d = cadd(d, cscale(h[c], fv[q][c]));
      acc[q] = cadd(acc[q], cscale(d, a.w[t]));
    }
  }
  const double A = a.area[f];
  for (int q = 0; q < 3; q++) a.side_out[3 * f + q] = cscale(acc[q], A);
}

__global__ void nls_gather2_kernel(int E, const int *__restrict__ eside, const double2 *__restrict__ side_out,
                                   double *B) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double2 m = make_double2(0, 0);
  for (int k = 0; k < 2; k++) {
    const int s = eside[2 * e + k];
    if (s >= 0) m = cadd(m, side_out[s]);
  }
  B[e] = m.x;
  B[E + e] = m.y;
}

__global__ void negate_kernel(int64_t n, double *v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = -v[i];
}

// src_coef = (srcM, srcJ) from the two solves, src_vec = (0.5 coef, srcH)
__global__ void nls_output_kernel(int E, const double *__restrict__ X1, const double *__restrict__ X2,
                                  const double2 *__restrict__ coef0, const double2 *__restrict__ srch,
                                  double2 *src_coef, double2 *src_vec) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  src_coef[e] = make_double2(X2[e], X2[E + e]);
  src_coef[E + e] = make_double2(X1[2 * E + e], X1[3 * E + e]);
  src_vec[e] = cscale(coef0[e], 0.5);
  src_vec[E + e] = srch[e];
}

// Gram entries: up to two (face, side q, side p) terms per entry,
// value = sum_terms A sum_r w_r f_q(r) . f_p(r)  (rwgf.f90:167-218)
struct GramArgs {
  int64_t nnz;
  int Q;
  double w[kMaxQ];
  const double *qp, *area, *coef, *v;
  const int *term;  // [nnz][2][3] (face, q, p), face -1: unused
  double *val;
};

__global__ void gram_values_kernel(GramArgs a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= a.nnz) return;
  double v = 0;
  for (int t = 0; t < 2; t++) {
    const int *tm = a.term + 6 * k + 3 * t;
    const int f = tm[0];
    if (f < 0) continue;
    const int sq = 3 * f + tm[1], sp = 3 * f + tm[2];
    const double cq = a.coef[sq], cp = a.coef[sp];
    double in = 0;
    for (int r = 0; r < a.Q; r++) {
      const double *pt = a.qp + ((int64_t)f * a.Q + r) * 3;
      double d = 0;
      for (int c = 0; c < 3; c++) d += (cq * (pt[c] - a.v[3 * sq + c])) * (cp * (pt[c] - a.v[3 * sp + c]));
      in += a.w[r] * d;
    }
    v += in * a.area[f];
  }
  a.val[k] = v;
}

// Jacobi-preconditioned conjugate gradients for F X = B, one workgroup per
// real right-hand side column (B, X: [ncol][E]; W: [ncol][3][E] scratch).
// Every reduction sums the waves' partials in one fixed order, so the
// iteration is deterministic.  Ends when the recursive residual falls to
// kCgTol of |b| (2-norms) or after maxit steps; info[col] = (steps, |r|/|b|).
struct CgArgs {
  int E, maxit;
  const int *rowp, *col;
  const double *val, *dinv, *B;
  double *X, *W, *info;
};

__device__ double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int k = 0; k < kCgThreads / 64; k++) s += red[k];
  return s;
}

__global__ __launch_bounds__(kCgThreads) void gram_cg_kernel(CgArgs a) {
  __shared__ double red[kCgThreads / 64];
  const int E = a.E, c = blockIdx.x;
  const double *b = a.B + (int64_t)c * E;
  double *x = a.X + (int64_t)c * E, *r = a.W + (int64_t)c * 3 * E, *p = r + E, *q = p + E;
  double bb = 0, rz = 0;
  for (int i = threadIdx.x; i < E; i += kCgThreads) {
    const double bi = b[i], zi = bi * a.dinv[i];
    x[i] = 0;
    r[i] = bi;
    p[i] = zi;
    bb += bi * bi;
    rz += bi * zi;
  }
  bb = block_sum(bb, red);
  rz = block_sum(rz, red);
  int it = 0;
  double rr = bb;
  if (bb > 0) {
    const double stop = kCgTol * kCgTol * bb;
    for (it = 1; it <= a.maxit; it++) {
      double pq = 0;
      for (int i = threadIdx.x; i < E; i += kCgThreads) {
        double s = 0;
        for (int k = a.rowp[i]; k < a.rowp[i + 1]; k++) s += a.val[k] * p[a.col[k]];
        q[i] = s;
        pq += p[i] * s;
      }
      pq = block_sum(pq, red);
      const double alpha = rz / pq;
      double rrl = 0, rzl = 0;
      for (int i = threadIdx.x; i < E; i += kCgThreads) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        rrl += ri * ri;
        rzl += ri * (ri * a.dinv[i]);
      }
      rr = block_sum(rrl, red);
      const double rzn = block_sum(rzl, red);
      if (!(rr > stop)) break;
      const double beta = rzn / rz;
      rz = rzn;
      for (int i = threadIdx.x; i < E; i += kCgThreads) p[i] = r[i] * a.dinv[i] + beta * p[i];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    a.info[2 * c] = (double)it;
    a.info[2 * c + 1] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  }
}

__global__ void gram_dinv_kernel(int E, const int *__restrict__ rowp, const int *__restrict__ col,
                                 const double *__restrict__ val, double *dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E) return;
  double d = 0;
  for (int k = rowp[i]; k < rowp[i + 1]; k++)
    if (col[k] == i) d = val[k];
  dinv[i] = d > 0 ? 1.0 / d : 0.0;
}

// Sparse Gram matrix of one mesh on the device (CSR, columns ascending).
struct Gram {
  int E = 0;
  int64_t nnz = 0;
  DevBuf rowp, col, val, dinv;
};

static int local_side(const mom_mesh *m, int f, int e) {
  for (int q = 0; q < 3; q++)
    if (m->h_fedges[3 * f + q] == e) return q;
  return -1;
}

static int build_gram(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, Gram &G) {
  const int E = (int)m->E;
  struct Ent {
    int col, t[6];
  };
  std::vector<std::vector<Ent>> rows(E);
  for (int a = 0; a < E; a++) {
    Ent d{a, {-1, 0, 0, -1, 0, 0}};
    int nt = 0;
    for (int s = 0; s < 2; s++) {
      const int f = m->h_efaces[2 * a + s];
      if (f < 0) continue;
      const int q = local_side(m, f, a);
      if (q < 0) return fail(MOM_ERR_ARG, "Invalid edge index request (nlsurf.f90)!");
      d.t[3 * nt] = f;
      d.t[3 * nt + 1] = q;
      d.t[3 * nt + 2] = q;
      nt++;
    }
    rows[a].push_back(d);
  }
  // off-diagonal entries on the first common face of the lower edge
  // (rwgf.f90:191-201)
  for (int f = 0; f < (int)m->F; f++)
    for (int i = 0; i < 3; i++)
      for (int j = i + 1; j < 3; j++) {
        const int ea = std::min(m->h_fedges[3 * f + i], m->h_fedges[3 * f + j]);
        const int eb = std::max(m->h_fedges[3 * f + i], m->h_fedges[3 * f + j]);
        if (ea == eb) continue;
        const int a0 = m->h_efaces[2 * ea], a1 = m->h_efaces[2 * ea + 1];
        const int b0 = m->h_efaces[2 * eb], b1 = m->h_efaces[2 * eb + 1];
        int fc = -1;
        if (a0 != -1 && (a0 == b0 || a0 == b1))
          fc = a0;
        else if (a1 != -1 && (a1 == b0 || a1 == b1))
          fc = a1;
        if (fc != f) continue;
        const int q = local_side(m, f, ea), p = local_side(m, f, eb);
        rows[ea].push_back(Ent{eb, {f, q, p, -1, 0, 0}});
        rows[eb].push_back(Ent{ea, {f, p, q, -1, 0, 0}});
      }
  std::vector<int> rowp(E + 1, 0), col;
  std::vector<int> term;
  for (int a = 0; a < E; a++) {
    auto &r = rows[a];
    std::sort(r.begin(), r.end(), [](const Ent &x, const Ent &y) { return x.col < y.col; });
    for (const Ent &e : r) {
      col.push_back(e.col);
      term.insert(term.end(), e.t, e.t + 6);
    }
    rowp[a + 1] = (int)col.size();
  }
  G.E = E;
  G.nnz = (int64_t)col.size();
  DevBuf dterm;
  int rc;
  if ((rc = G.rowp.alloc(sizeof(int) * (E + 1))) || (rc = G.col.alloc(sizeof(int) * G.nnz)) ||
      (rc = G.val.alloc(sizeof(double) * G.nnz)) || (rc = G.dinv.alloc(sizeof(double) * E)) ||
      (rc = dterm.alloc(sizeof(int) * term.size())))
    return rc;
  MOM_TRY(copy_h2d(ctx, G.rowp.ptr, rowp.data(), sizeof(int) * (E + 1)));
  MOM_TRY(copy_h2d(ctx, G.col.ptr, col.data(), sizeof(int) * G.nnz));
  MOM_TRY(copy_h2d(ctx, dterm.ptr, term.data(), sizeof(int) * term.size()));
  GramArgs ga;
  ga.nnz = G.nnz;
  ga.Q = rule.Q;
  for (int i = 0; i < kMaxQ; i++) ga.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  ga.qp = ctx->qp.as<double>();
  ga.area = m->farea.as<double>();
  ga.coef = m->s_coef.as<double>();
  ga.v = m->s_v.as<double>();
  ga.term = dterm.as<int>();
  ga.val = G.val.as<double>();
  hipLaunchKernelGGL(gram_values_kernel, dim3((unsigned)((G.nnz + 255) / 256)), dim3(256), 0, ctx->stream, ga);
  MOM_HIP(hipGetLastError());
  hipLaunchKernelGGL(gram_dinv_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     G.rowp.as<int>(), G.col.as<int>(), G.val.as<double>(), G.dinv.as<double>());
  MOM_HIP(hipGetLastError());
  // the term table must outlive the kernel
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int gram_solve(mom_ctx *ctx, const Gram &G, int ncol, const double *B, double *X, double *W, double *info_d) {
  CgArgs c;
  c.E = G.E;
  c.maxit = std::min(G.E + 100, 20000);
  c.rowp = G.rowp.as<int>();
  c.col = G.col.as<int>();
  c.val = G.val.as<double>();
  c.dinv = G.dinv.as<double>();
  c.B = B;
  c.X = X;
  c.W = W;
  c.info = info_d;
  hipLaunchKernelGGL(gram_cg_kernel, dim3((unsigned)ncol), dim3(kCgThreads), 0, ctx->stream, c);
  MOM_HIP(hipGetLastError());
  std::vector<double> info(2 * ncol);
  MOM_TRY(copy_d2h(ctx, info.data(), info_d, sizeof(double) * 2 * ncol));
  for (int k = 0; k < ncol; k++)
    if (!(info[2 * k + 1] <= 1e-12))
      return fail(MOM_ERR_NOCONV, "Gram solve (RWG moments) did not converge: relative residual " +
                                      std::to_string(info[2 * k + 1]) + " after " +
                                      std::to_string((int)info[2 * k]) + " steps");
  return 0;
}

static int nlsurf_coef_impl(mom_ctx *ctx, const mom_mesh *m, int64_t Nt, double omega, cd riff,
                            const double *epsp, const double *x, const double *chi2, const QuadRule &rule,
                            double *src_coef, double *src_vec) {
  if (m->max_parent >= Nt) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  for (int64_t f = 0; f < m->F; f++)
    if (!(epsp[2 * f] != 0 || epsp[2 * f + 1] != 0)) return fail(MOM_ERR_ARG, "epsp must be non-zero");
  MOM_TRY(compute_quad_points(ctx, m, rule));
  const int F = (int)m->F, E = (int)m->E;
  DevBuf dx, deps, dside, dB1, dX1, dB2, dX2, dW, dinfo, dcoef0, dsrch, dsc, dsv, deside;
  int rc;
  if ((rc = dx.alloc(sizeof(double2) * 2 * Nt)) || (rc = deps.alloc(sizeof(double2) * F)) ||
      (rc = dside.alloc(sizeof(double2) * 9 * F)) || (rc = dB1.alloc(sizeof(double) * 4 * E)) ||
      (rc = dX1.alloc(sizeof(double) * 4 * E)) || (rc = dB2.alloc(sizeof(double) * 2 * E)) ||
      (rc = dX2.alloc(sizeof(double) * 2 * E)) || (rc = dW.alloc(sizeof(double) * 12 * E)) ||
      (rc = dinfo.alloc(sizeof(double) * 8)) || (rc = dcoef0.alloc(sizeof(double2) * E)) ||
      (rc = dsrch.alloc(sizeof(double2) * E)) || (rc = dsc.alloc(sizeof(double2) * 2 * E)) ||
      (rc = dsv.alloc(sizeof(double2) * 2 * E)) || (rc = deside.alloc(sizeof(int) * 2 * E)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dx.ptr, x, sizeof(double2) * 2 * Nt));
  MOM_TRY(copy_h2d(ctx, deps.ptr, epsp, sizeof(double2) * F));
  MOM_TRY(copy_h2d(ctx, deside.ptr, m->h_eside.data(), sizeof(int) * 2 * E));
  const cd I(0.0, 1.0);
  const cd enf = 1.0 / ((I * omega) * (kEps0 * (riff * riff)));
  NlsFaceArgs a;
  a.F = F;
  a.Q = rule.Q;
  for (int i = 0; i < kMaxQ; i++) a.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  a.qp = ctx->qp.as<double>();
  a.fn = m->fn.as<double>();
  a.area = m->farea.as<double>();
  a.g = m->s_g.as<int>();
  a.coef = m->s_coef.as<double>();
  a.div = m->s_div.as<double>();
  a.v = m->s_v.as<double>();
  a.x = dx.as<double2>();
  a.Nt = Nt;
  a.en_fac = make_double2(enf.real(), enf.imag());
  for (int k = 0; k < 3; k++) a.chi[k] = make_double2(chi2[2 * k], chi2[2 * k + 1]);
  a.epsp = deps.as<double2>();
  a.side_out = dside.as<double2>();
  const double omegash = 2.0 * omega;
  hipLaunchKernelGGL(nls_stage1_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, ctx->stream, a);
  MOM_HIP(hipGetLastError());
  hipLaunchKernelGGL(nls_gather1_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     deside.as<int>(), dside.as<double2>(), omegash, dB1.as<double>(), dcoef0.as<double2>(),
                     dsrch.as<double2>());
  MOM_HIP(hipGetLastError());
  Gram G;
  MOM_TRY(build_gram(ctx, m, rule, G));
  MOM_TRY(gram_solve(ctx, G, 4, dB1.as<double>(), dX1.as<double>(), dW.as<double>(), dinfo.as<double>()));
  NlsStage2Args b;
  b.F = F;
  b.Q = rule.Q;
  b.E = E;
  for (int i = 0; i < kMaxQ; i++) b.w[i] = a.w[i];
  b.qp = a.qp;
  b.fn = a.fn;
  b.area = a.area;
  b.l = m->s_l.as<int>();
  b.coef = a.coef;
  b.v = a.v;
  b.X = dX1.as<double>();
  b.side_out = dside.as<double2>();
  hipLaunchKernelGGL(nls_stage2_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, ctx->stream, b);
  MOM_HIP(hipGetLastError());
  hipLaunchKernelGGL(nls_gather2_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     deside.as<int>(), dside.as<double2>(), dB2.as<double>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(gram_solve(ctx, G, 2, dB2.as<double>(), dX2.as<double>(), dW.as<double>(), dinfo.as<double>()));
  hipLaunchKernelGGL(nls_output_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     dX1.as<double>(), dX2.as<double>(), dcoef0.as<double2>(), dsrch.as<double2>(),
                     dsc.as<double2>(), dsv.as<double2>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(copy_d2h(ctx, src_coef, dsc.ptr, sizeof(double2) * 2 * E));
  MOM_TRY(copy_d2h(ctx, src_vec, dsv.ptr, sizeof(double2) * 2 * E));
  return 0;
}

// ---------------------------------------------------------------------------
// Adjoint of nlsurf_coef: the outputs src_coef = (srcM, srcJ) and
// src_vec = (coef0 / 2, srcH) are holomorphic and quadratic in the
// fundamental solution x, so for adjoint values (scb, svb) of the outputs
// the transpose of the Jacobian at x gives xb with
// Re(xb^T dx) = Re(scb^T dsrc_coef + svb^T dsrc_vec).  Backwards through
// the forward steps:
//   srcM = G^-1 m, srcJ = G^-1 jj:  mb = G^-1 srcMb, jjb = G^-1 srcJb
//     (G real symmetric: its own transpose);
//   m = gather(stage2(coef)): the per-face map of stage2 is antisymmetric
//     ((n x a) . b = -(n x b) . a), so coefb = -gather(stage2(mb));
//   coef = G^-1 c:  cb = G^-1 coefb + svb_J / 2;
//   jj = -i omegash j:  jb = -i omegash jjb;  srcH = i omegash h / 2:
//     hb = i omegash svb_M / 2;
//   stage 1 (per face and point) reversed by nls_stage1_adj_kernel with the
//   field and polarisation of x recomputed.
// Per edge (c, j, h) adjoints: eb[3][E] = (cb, jb, hb).
__global__ void nls_adj_edges_kernel(int E, const double *__restrict__ Xa, const double *__restrict__ Xc,
                                     const double2 *__restrict__ svb, double omegash, double2 *__restrict__ eb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double2 s1 = svb[e], s2 = svb[E + e];
  eb[e] = make_double2(Xc[e] + 0.5 * s1.x, Xc[E + e] + 0.5 * s1.y);
  // -i omegash (a + i b) = omegash (b - i a)
  eb[E + e] = make_double2(omegash * Xa[3 * E + e], -omegash * Xa[2 * E + e]);
  // i omegash / 2 (a + i b) = omegash / 2 (-b + i a)
  eb[2 * E + e] = make_double2(-0.5 * omegash * s2.y, 0.5 * omegash * s2.x);
}

// Stage 1 reversed, one thread per face: side q's output adjoints
// (i1, i3, i2 scaled as the forward writes them) come from its edge's eb;
// w[2 (3f + q)] = d/dxJ_q, w[2 (3f + q) + 1] = d/dxM_q (transposed).
__global__ __launch_bounds__(128) void nls_stage1_adj_kernel(NlsFaceArgs a, const int *__restrict__ s_l, int E,
                                                             const double2 *__restrict__ eb, double2 *__restrict__ w) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const double n[3] = {a.fn[3 * f], a.fn[3 * f + 1], a.fn[3 * f + 2]};
  const double A = a.area[f];
  const double2 ep = a.epsp[f];
  double2 xj[3], xm[3], b1[3], b2[3], b3[3];
  double cf[3], dv[3], vv[3][3];
  for (int q = 0; q < 3; q++) {
    const int s = 3 * f + q;
    xj[q] = a.x[a.g[s]];
    xm[q] = a.x[a.Nt + a.g[s]];
    cf[q] = a.coef[s];
    dv[q] = a.div[s];
    for (int c = 0; c < 3; c++) vv[q][c] = a.v[3 * s + c];
    const int l = s_l[s];
    const double2 z = make_double2(0, 0);
    // out = (i1 A / ep, i3 A, i2 A)
    b1[q] = l >= 0 ? cdiv(cscale(eb[l], A), ep) : z;
    b3[q] = l >= 0 ? cscale(eb[E + l], A) : z;
    b2[q] = l >= 0 ? cscale(eb[2 * E + l], A) : z;
  }
  double2 divj = make_double2(0, 0);
  for (int q = 0; q < 3; q++) divj = cadd(divj, cscale(xj[q], dv[q]));
  const double2 en = cmul(divj, a.en_fac);
  double2 xmb[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  double2 enb = make_double2(0, 0);
  for (int t = 0; t < a.Q; t++) {
    const double *r = a.qp + ((int64_t)f * a.Q + t) * 3;
    double fv[3][3];
    for (int q = 0; q < 3; q++)
      for (int c = 0; c < 3; c++) fv[q][c] = cf[q] * (r[c] - vv[q][c]);
    // forward values at this point
    double2 ms[3];
    for (int c = 0; c < 3; c++) {
      ms[c] = make_double2(0, 0);
      for (int q = 0; q < 3; q++) ms[c] = cadd(ms[c], cscale(xm[q], fv[q][c]));
    }
    double2 e[3];
    e[0] = cadd(csub(cscale(ms[2], n[1]), cscale(ms[1], n[2])), cscale(en, n[0]));
    e[1] = cadd(csub(cscale(ms[0], n[2]), cscale(ms[2], n[0])), cscale(en, n[1]));
    e[2] = cadd(csub(cscale(ms[1], n[0]), cscale(ms[0], n[1])), cscale(en, n[2]));
    const double2 enn = cadd(cadd(cscale(e[0], n[0]), cscale(e[1], n[1])), cscale(e[2], n[2]));
    double2 et[3];
    for (int c = 0; c < 3; c++) et[c] = csub(e[c], cscale(enn, n[c]));
    const double2 tfac = cmul(cscale(a.chi[2], 2.0 * kEps0), enn);
    // adjoints of pn, ptn and pt from i1, i2, i3
    const double wt = a.w[t];
    double2 pnb = make_double2(0, 0), ptnb[3], ptb[3];
    for (int c = 0; c < 3; c++) ptnb[c] = ptb[c] = make_double2(0, 0);
    for (int q = 0; q < 3; q++) {
      pnb = csub(pnb, cscale(b1[q], wt * dv[q]));
      for (int c = 0; c < 3; c++) {
        ptnb[c] = cadd(ptnb[c], cscale(b2[q], wt * fv[q][c]));
        ptb[c] = cadd(ptb[c], cscale(b3[q], wt * fv[q][c]));
      }
    }
    // ptn = pt x n:  ptb += n x ptnb
    ptb[0] = cadd(ptb[0], csub(cscale(ptnb[2], n[1]), cscale(ptnb[1], n[2])));
    ptb[1] = cadd(ptb[1], csub(cscale(ptnb[0], n[2]), cscale(ptnb[2], n[0])));
    ptb[2] = cadd(ptb[2], csub(cscale(ptnb[1], n[0]), cscale(ptnb[0], n[1])));
    // pt = P - pn n, pn = n . P:  Pb = ptb + n (pnb - n . ptb)
    const double2 nptb = cadd(cadd(cscale(ptb[0], n[0]), cscale(ptb[1], n[1])), cscale(ptb[2], n[2]));
    const double2 pnt = csub(pnb, nptb);
    double2 Pb[3];
    for (int c = 0; c < 3; c++) Pb[c] = cadd(ptb[c], cscale(pnt, n[c]));
    // P = et tfac + pn0 n
    double2 etb[3];
    for (int c = 0; c < 3; c++) etb[c] = cmul(tfac, Pb[c]);
    const double2 tfacb = cadd(cadd(cmul(et[0], Pb[0]), cmul(et[1], Pb[1])), cmul(et[2], Pb[2]));
    const double2 pn0b = cadd(cadd(cscale(Pb[0], n[0]), cscale(Pb[1], n[1])), cscale(Pb[2], n[2]));
    // tfac = 2 eps0 chi_ttn enn; pn0 = eps0 (chi_nnn enn^2 + chi_ntt ett)
    double2 ennb = cmul(cscale(a.chi[2], 2.0 * kEps0), tfacb);
    ennb = cadd(ennb, cmul(cscale(cmul(a.chi[0], enn), 2.0 * kEps0), pn0b));
    const double2 ettb = cmul(cscale(a.chi[1], kEps0), pn0b);
    // ett = et . et
    for (int c = 0; c < 3; c++) etb[c] = cadd(etb[c], cmul(cscale(et[c], 2.0), ettb));
    // et = e - enn n, enn = n . e:  eb = etb + n (ennb - n . etb)
    const double2 netb = cadd(cadd(cscale(etb[0], n[0]), cscale(etb[1], n[1])), cscale(etb[2], n[2]));
    const double2 ent = csub(ennb, netb);
    double2 ebv[3];
    for (int c = 0; c < 3; c++) ebv[c] = cadd(etb[c], cscale(ent, n[c]));
    // e = n x ms + en n:  msb = ebv x n, enb += n . ebv
    double2 msb[3];
    msb[0] = csub(cscale(ebv[1], n[2]), cscale(ebv[2], n[1]));
    msb[1] = csub(cscale(ebv[2], n[0]), cscale(ebv[0], n[2]));
    msb[2] = csub(cscale(ebv[0], n[1]), cscale(ebv[1], n[0]));
    enb = cadd(enb, cadd(cadd(cscale(ebv[0], n[0]), cscale(ebv[1], n[1])), cscale(ebv[2], n[2])));
    // ms = sum_q xm_q f_q
    for (int q = 0; q < 3; q++)
      for (int c = 0; c < 3; c++) xmb[q] = cadd(xmb[q], cscale(msb[c], fv[q][c]));
  }
  // en = divj en_fac, divj = sum_q xj_q div_q
  const double2 divjb = cmul(enb, a.en_fac);
  for (int q = 0; q < 3; q++) {
    w[2 * (3 * f + q)] = cscale(divjb, dv[q]);
    w[2 * (3 * f + q) + 1] = xmb[q];
  }
}

static int nlsurf_coef_adjoint_impl(mom_ctx *ctx, const mom_mesh *m, int64_t Nt, double omega, cd riff,
                                    const double *epsp, const double *x, const double *chi2, const QuadRule &rule,
                                    const double *scb, const double *svb, double *xb) {
  if (m->max_parent >= Nt) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  for (int64_t f = 0; f < m->F; f++)
    if (!(epsp[2 * f] != 0 || epsp[2 * f + 1] != 0)) return fail(MOM_ERR_ARG, "epsp must be non-zero");
  MOM_TRY(compute_quad_points(ctx, m, rule));
  const int F = (int)m->F, E = (int)m->E;
  DevBuf dx, deps, dside, dBa, dXa, dB2, dXc, dW, dinfo, dsvb, deb, dw, dg, deside;
  int rc;
  if ((rc = dx.alloc(sizeof(double2) * 2 * Nt)) || (rc = deps.alloc(sizeof(double2) * F)) ||
      (rc = dside.alloc(sizeof(double2) * 3 * F)) || (rc = dBa.alloc(sizeof(double) * 4 * E)) ||
      (rc = dXa.alloc(sizeof(double) * 4 * E)) || (rc = dB2.alloc(sizeof(double) * 2 * E)) ||
      (rc = dXc.alloc(sizeof(double) * 2 * E)) || (rc = dW.alloc(sizeof(double) * 12 * E)) ||
      (rc = dinfo.alloc(sizeof(double) * 8)) || (rc = dsvb.alloc(sizeof(double2) * 2 * E)) ||
      (rc = deb.alloc(sizeof(double2) * 3 * E)) || (rc = dw.alloc(sizeof(double2) * 6 * F)) ||
      (rc = dg.alloc(sizeof(double2) * 2 * Nt)) || (rc = deside.alloc(sizeof(int) * 2 * E)))
    return rc;
  // right-hand sides of the first adjoint Gram solve: Re, Im of srcMb, then of srcJb
  std::vector<double> Ba(4 * (size_t)E);
  for (int e = 0; e < E; e++) {
    Ba[e] = scb[2 * e];
    Ba[E + e] = scb[2 * e + 1];
    Ba[2 * E + e] = scb[2 * (E + e)];
    Ba[3 * E + e] = scb[2 * (E + e) + 1];
  }
  MOM_TRY(copy_h2d(ctx, dx.ptr, x, sizeof(double2) * 2 * Nt));
  MOM_TRY(copy_h2d(ctx, deps.ptr, epsp, sizeof(double2) * F));
  MOM_TRY(copy_h2d(ctx, dBa.ptr, Ba.data(), sizeof(double) * 4 * E));
  MOM_TRY(copy_h2d(ctx, dsvb.ptr, svb, sizeof(double2) * 2 * E));
  MOM_HIP(hipMemsetAsync(dg.ptr, 0, dg.bytes, ctx->stream));
  Gram G;
  MOM_TRY(build_gram(ctx, m, rule, G));
  MOM_TRY(gram_solve(ctx, G, 4, dBa.as<double>(), dXa.as<double>(), dW.as<double>(), dinfo.as<double>()));
  // coefb = -gather(stage2(mb)), mb = rows 0, 1 of Xa
  NlsStage2Args b;
  b.F = F;
  b.Q = rule.Q;
  b.E = E;
  for (int i = 0; i < kMaxQ; i++) b.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  b.qp = ctx->qp.as<double>();
  b.fn = m->fn.as<double>();
  b.area = m->farea.as<double>();
  b.l = m->s_l.as<int>();
  b.coef = m->s_coef.as<double>();
  b.v = m->s_v.as<double>();
  b.X = dXa.as<double>();
  b.side_out = dside.as<double2>();
  MOM_TRY(copy_h2d(ctx, deside.ptr, m->h_eside.data(), sizeof(int) * 2 * E));
  hipLaunchKernelGGL(nls_stage2_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, ctx->stream, b);
  hipLaunchKernelGGL(nls_gather2_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     deside.as<int>(), dside.as<double2>(), dB2.as<double>());
  hipLaunchKernelGGL(negate_kernel, dim3((unsigned)((2 * E + 255) / 256)), dim3(256), 0, ctx->stream,
                     (int64_t)2 * E, dB2.as<double>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(gram_solve(ctx, G, 2, dB2.as<double>(), dXc.as<double>(), dW.as<double>(), dinfo.as<double>()));
  const double omegash = 2.0 * omega;
  hipLaunchKernelGGL(nls_adj_edges_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, ctx->stream, E,
                     dXa.as<double>(), dXc.as<double>(), dsvb.as<double2>(), omegash, deb.as<double2>());
  const cd I(0.0, 1.0);
  const cd enf = 1.0 / ((I * omega) * (kEps0 * (riff * riff)));
  NlsFaceArgs a;
  a.F = F;
  a.Q = rule.Q;
  for (int i = 0; i < kMaxQ; i++) a.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  a.qp = ctx->qp.as<double>();
  a.fn = m->fn.as<double>();
  a.area = m->farea.as<double>();
  a.g = m->s_g.as<int>();
  a.coef = m->s_coef.as<double>();
  a.div = m->s_div.as<double>();
  a.v = m->s_v.as<double>();
  a.x = dx.as<double2>();
  a.Nt = Nt;
  a.en_fac = make_double2(enf.real(), enf.imag());
  for (int k = 0; k < 3; k++) a.chi[k] = make_double2(chi2[2 * k], chi2[2 * k + 1]);
  a.epsp = deps.as<double2>();
  a.side_out = nullptr;
  hipLaunchKernelGGL(nls_stage1_adj_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, ctx->stream, a,
                     m->s_l.as<int>(), E, deb.as<double2>(), dw.as<double2>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(gather_sides(ctx, m, dw.as<double2>(), Nt, dg.as<double2>()));
  return copy_d2h(ctx, xb, dg.ptr, sizeof(double2) * 2 * Nt);
}

// Z[g(i) + bi Nt, g(j) + bj Nt] += Zd[i + bi E, j + bj E]: one domain's local
// PMCHWT block added into the global matrix (g injective: no conflicts)
__global__ void scatter_block_kernel(const double2 *__restrict__ Zd, int64_t E, const int *__restrict__ g,
                                     int64_t Nt, double2 *Z, int64_t ldz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 2 * E) return;
  const int64_t gi = g[i % E] + (i >= E ? Nt : 0);
  for (int64_t j = blockIdx.y; j < 2 * E; j += gridDim.y) {
    const int64_t gj = g[j % E] + (j >= E ? Nt : 0);
    const double2 v = Zd[i + j * 2 * E];
    double2 *z = Z + gi + gj * ldz;
    const double2 o = *z;
    *z = make_double2(o.x + v.x, o.y + v.y);
  }
}

int fill_system_nls(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *shri,
                    int64_t nedges_total, double omega, int32_t nq, const double *qw, const double *qnodes,
                    double2 *Z, int64_t ldz, int64_t nsrc, const double *src_coef, int64_t ldsc, double *src,
                    int64_t ldsrc) {
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  if (nsrc < 0 || nsrc > 65535 || (nsrc > 0 && (!src_coef || !src))) return fail(MOM_ERR_ARG, "bad source arrays");
  const int64_t Nt = nedges_total, N2 = 2 * Nt;
  if (nsrc > 0 && (ldsrc < N2 || ldsc < 1)) return fail(MOM_ERR_ARG, "ldsrc < 2N or ldsc < 1");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  for (int d = 0; d < ndom; d++) {
    if (!doms[d]) return fail(MOM_ERR_ARG, "NULL domain mesh");
    if (doms[d]->ctx != ctx) return fail(MOM_ERR_ARG, "domain mesh belongs to another context");
    if (doms[d]->max_parent >= Nt) return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
    if (nsrc > 0 && ldsc < 2 * doms[d]->E) return fail(MOM_ERR_ARG, "ldsc < 2 * domain edges");
  }
  MOM_HIP(hipMemset2DAsync(Z, ldz * sizeof(double2), 0, N2 * sizeof(double2), N2, ctx->stream));
  for (int d = 0; d < ndom; d++) {
    const mom_mesh *m = doms[d];
    const cd r(shri[2 * d], shri[2 * d + 1]);
    const cd zsi = (r * r) / (kEta0 * kEta0);
    const int64_t E = m->E;
    // src_coef column of (domain d, source s): src_coef + 2 ldsc (d + ndom s)
    bool active = false;
    for (int64_t s = 0; s < nsrc && !active; s++) {
      const double *c = src_coef + 2 * ldsc * (d + (int64_t)ndom * s);
      for (int64_t i = 0; i < 4 * E; i++)
        if (c[i] != 0) {
          active = true;
          break;
        }
    }
    FillTarget t;
    t.mode = 0;
    t.zsi = make_double2(zsi.real(), zsi.imag());
    if (!active) {
      t.Z = Z;
      t.ld = ldz;
      t.Noff = Nt;
      MOM_TRY(fill_domain(ctx, m, rule, omega, r, t));
      continue;
    }
    DevBuf Zd, dV, dY, dg;
    if ((rc = Zd.alloc(sizeof(double2) * 4 * E * E)) || (rc = dV.alloc(sizeof(double2) * 2 * E * nsrc)) ||
        (rc = dY.alloc(sizeof(double2) * 2 * E * nsrc)) || (rc = dg.alloc(sizeof(int) * E)))
      return rc;
    MOM_HIP(hipMemsetAsync(Zd.ptr, 0, Zd.bytes, ctx->stream));
    t.Z = Zd.as<double2>();
    t.ld = 2 * E;
    t.Noff = E;
    t.local = 1;
    MOM_TRY(fill_domain(ctx, m, rule, omega, r, t));
    // v = (srcJ, srcM) per source: src_coef holds (srcM, srcJ)
    std::vector<double> V(4 * E * nsrc), Y(4 * E * nsrc);
    for (int64_t s = 0; s < nsrc; s++) {
      const double *c = src_coef + 2 * ldsc * (d + (int64_t)ndom * s);
      double *v = V.data() + 4 * E * s;
      std::copy(c + 2 * E, c + 4 * E, v);
      std::copy(c, c + 2 * E, v + 2 * E);
    }
    MOM_TRY(copy_h2d(ctx, dV.ptr, V.data(), sizeof(double) * V.size()));
    MOM_TRY(copy_h2d(ctx, dg.ptr, m->h_eparent.data(), sizeof(int) * E));
    for (int64_t sidx = 0; sidx < nsrc; sidx++)
      MOM_TRY(matvec(ctx, Zd.as<double2>(), 2 * E, 2 * E, 2 * E, dV.as<double2>() + 2 * E * sidx,
                     dY.as<double2>() + 2 * E * sidx));
    hipLaunchKernelGGL(scatter_block_kernel, dim3((unsigned)((2 * E + 255) / 256), (unsigned)std::min<int64_t>(2 * E, 32768)), dim3(256), 0,
                       ctx->stream, Zd.as<double2>(), E, dg.as<int>(), Nt, Z, ldz);
    MOM_HIP(hipGetLastError());
    MOM_TRY(copy_d2h(ctx, Y.data(), dY.ptr, sizeof(double) * Y.size()));
    for (int64_t s = 0; s < nsrc; s++) {
      const double *y = Y.data() + 4 * E * s;
      double *b = src + 2 * ldsrc * s;
      for (int64_t i = 0; i < E; i++) {
        const int64_t gi = m->h_eparent[i];
        b[2 * gi] -= y[2 * i];
        b[2 * gi + 1] -= y[2 * i + 1];
        b[2 * (gi + Nt)] -= y[2 * (E + i)];
        b[2 * (gi + Nt) + 1] -= y[2 * (E + i) + 1];
      }
    }
  }
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace mom

using namespace mom;

extern "C" {

int mom_nlsurf_coef(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                    double ri_im, const double *epsp, const double *x, const double *chi2, int32_t nq,
                    const double *qw, const double *qnodes, double *src_coef, double *src_vec) {
  if (!ctx || !mesh || !epsp || !x || !chi2 || !src_coef || !src_vec) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  (void)hipSetDevice(ctx->device);
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  return nlsurf_coef_impl(ctx, mesh, nedges_total, omega, cd(ri_re, ri_im), epsp, x, chi2, rule, src_coef,
                          src_vec);
}

int mom_nlsurf_coef_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                            double ri_im, const double *epsp, const double *x, const double *chi2, int32_t nq,
                            const double *qw, const double *qnodes, const double *src_coef_bar,
                            const double *src_vec_bar, double *xbar) {
  if (!ctx || !mesh || !epsp || !x || !chi2 || !src_coef_bar || !src_vec_bar || !xbar)
    return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  (void)hipSetDevice(ctx->device);
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  return nlsurf_coef_adjoint_impl(ctx, mesh, nedges_total, omega, cd(ri_re, ri_im), epsp, x, chi2, rule,
                                  src_coef_bar, src_vec_bar, xbar);
}

int mom_sysmat_pmchwt_nls(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *shri,
                          int64_t nedges_total, double omega, int32_t nq, const double *qw, const double *qnodes,
                          int64_t nsrc, const double *src_coef, int64_t ldsc, double *src, int64_t ldsrc,
                          double *A, int64_t lda) {
  if (!ctx || !doms || !shri || !A) return fail(MOM_ERR_ARG, "NULL argument");
  if (ndom < 1 || nedges_total < 1) return fail(MOM_ERR_ARG, "bad sizes");
  const int64_t N2 = 2 * nedges_total;
  if (lda < N2) return fail(MOM_ERR_ARG, "lda < 2N");
  (void)hipSetDevice(ctx->device);
  DevBuf Z;
  MOM_TRY(Z.alloc(sizeof(double2) * N2 * N2));
  MOM_TRY(fill_system_nls(ctx, ndom, doms, shri, nedges_total, omega, nq, qw, qnodes, Z.as<double2>(), N2, nsrc,
                          src_coef, ldsc, src, ldsrc));
  return copy_d2h_2d(ctx, A, lda * sizeof(double2), Z.ptr, N2 * sizeof(double2), N2 * sizeof(double2), N2);
}

}  // extern "C"
