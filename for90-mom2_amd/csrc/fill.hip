// PMCHWT system fill on gfx950: far/smooth pair kernel, singular-correction
// kernel, plane-wave right-hand sides and surface cross-sections.
//
// Reference algorithm (for90-mom2): sysmat.f90:18-105 (assembly),
// 272-519 (computeD/H/K), srcint.f90:57-448 (intK1/2/4), green.f90:25-151
// (G, grad G, smoothed kernels), int.f90 (analytic singular integrals),
// quad.f90:352-367 (quadrature points), rwgf.f90:34-78 (RWG functions),
// source.f90:63-247 (plane-wave RHS), cs.f90:14-104 (cross-sections).
//
// Design.  The reference evaluates, for every test face m, test point r_i
// and source face n, the Green's function at the 13 source points three
// times (D and H each call G, K calls grad G) and then contracts with the RWG
// functions edge by edge.  Here one thread owns one face pair (m, n) and
// evaluates exp(ikR) once per point pair, feeding both kernels G and
// g = G (1/R^2 - ik/R) (grad' G = (r - r') g).  Because an RWG function is
// affine on its face, f(r) = coef (rho - nu) in the face's local coordinates
// (rho = r - centroid, nu = free vertex - centroid), every D and K entry of
// the 3x3 edge block of the pair is a fixed combination of 18 complex
// moments sum_ij w_i w_j {G, g} x {1, rho_i, rho'_j, rho_i.rho'_j,
// rho_i x rho'_j}; H needs only sum_ij w_i w_j G.  Local coordinates keep
// the moments free of cancellation between face-sized vectors.
//
// Work layout: a workgroup of 4 waves owns 64 source faces (one per lane)
// and 16 test faces (4 per wave, wave-uniform).  Source points sit in LDS
// as [point][xyz][lane] (conflict-free), test points are broadcast.  Near
// pairs (centroid distance < 3*avelen, sysmat.f90:233-260) use the smoothed
// kernels; their analytic singular parts are added by a second kernel that
// runs one thread per near pair over the pair list built on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

constexpr double kOofpi = 1.0 / (4.0 * MOM_PI);

struct FaceSide {
  const int *g;
  const double *coef, *div, *v;
};

struct FarArgs {
  int F, Q;
  double w[kMaxQ];
  const double *qp;   // [F][Q][3] absolute quadrature points
  const double *cp;   // [F][3]
  const double *ccls; // [F][3] centroids that classify near pairs (cp unless frozen to another mesh)
  const double *area; // [F]
  FaceSide side;      // per face side (3F)
  double kr, ki;      // wavenumber
  double thr2;        // near threshold^2
  double2 c1, c2;     // D and H prefactors
  double js[3];       // test faces reflected by diag(js) (all 1: no image)
  int image;          // an image (the self-pair rule does not apply)
  int gen;            // a non-diagonal image: test faces moved by M, near rule |c_m - J c_n|
  double M[9], J[9];  // M = J^T, row-major (gen)
  FillTarget out;
};

// exp(a) (cos b + i sin b) for the point-pair loop, where the library
// exp and sincos are about half of the instructions.  Cody-Waite reduction
// (Cody and Waite, "Software Manual for the Elementary Functions", 1980):
// a = n ln2 + r with ln2 in two parts (the first with 32 significant bits,
// so n * part exact for |n| < 2^21), |r| <= ln2/2; b = q pi/2 + t with pi/2
// in three parts of 33, 33 and 53 bits (q * part exact for |q| < 2^20),
// |t| <= pi/4.  Then Taylor polynomials in Horner form: exp(r) to degree 13
// (truncation below 5e-18), sin t to t^15 and cos t to t^16 (below 6e-17),
// coefficients 1/k! rounded to double.  Needs |b| < kCisMaxArg (the fill
// checks Re(k) times the mesh extent on the host); a is clamped to
// [-1000, 710]: exp underflows to 0 below and overflows to inf above.
constexpr double kCisMaxArg = 524288.0;  // 2^19

// cos b and sin b (the rotation part of cis_exp); cis_exp(a, b) with a = 0
// gives exactly these values, since its exponential is then 1 exactly.
__device__ __forceinline__ void cis(double b, double &c, double &s) {
  const double qf = rint(b * 0x1.45f306dc9c883p-1);  // b * 2/pi
  double t = fma(-qf, 0x1.921fb54400000p+0, b);
  t = fma(-qf, 0x1.0b4611a600000p-34, t);
  t = fma(-qf, 0x1.3198a2e037073p-69, t);
  const double t2 = t * t;
  double sp = -0x1.ae7f3e733b81fp-41;                 // -1/15!
  sp = fma(sp, t2, 0x1.6124613a86d09p-33);            // 1/13!
  sp = fma(sp, t2, -0x1.ae64567f544e4p-26);           // -1/11!
  sp = fma(sp, t2, 0x1.71de3a556c734p-19);            // 1/9!
  sp = fma(sp, t2, -0x1.a01a01a01a01ap-13);           // -1/7!
  sp = fma(sp, t2, 0x1.1111111111111p-7);             // 1/5!
  sp = fma(sp, t2, -0x1.5555555555555p-3);            // -1/3!
  const double sn = fma(sp * t2, t, t);
  double cp = 0x1.ae7f3e733b81fp-45;                  // 1/16!
  cp = fma(cp, t2, -0x1.93974a8c07c9dp-37);           // -1/14!
  cp = fma(cp, t2, 0x1.1eed8eff8d898p-29);            // 1/12!
  cp = fma(cp, t2, -0x1.27e4fb7789f5cp-22);           // -1/10!
  cp = fma(cp, t2, 0x1.a01a01a01a01ap-16);            // 1/8!
  cp = fma(cp, t2, -0x1.6c16c16c16c17p-10);           // -1/6!
  cp = fma(cp, t2, 0x1.5555555555555p-5);             // 1/4!
  cp = fma(cp, t2, -0.5);
  const double cs = fma(cp, t2, 1.0);
  const int q = (int)qf & 3;
  const double s1 = (q & 1) ? cs : sn, c1 = (q & 1) ? sn : cs;
  s = (q & 2) ? -s1 : s1;
  c = ((q + 1) & 2) ? -c1 : c1;
}

__device__ __forceinline__ void cis_exp(double a, double b, double &er, double &ei) {
  a = fmin(fmax(a, -1000.0), 710.0);
  // exp(a) = 2^n exp(r)
  const double nf = rint(a * 0x1.71547652b82fep+0);  // a / ln2
  double r = fma(-nf, 0x1.62e42fee00000p-1, a);
  r = fma(-nf, 0x1.a39ef35793c76p-33, r);
  double p = 0x1.6124613a86d09p-33;                   // 1/13!
  p = fma(p, r, 0x1.1eed8eff8d898p-29);               // 1/12!
  p = fma(p, r, 0x1.ae64567f544e4p-26);               // 1/11!
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);               // 1/10!
  p = fma(p, r, 0x1.71de3a556c734p-19);               // 1/9!
  p = fma(p, r, 0x1.a01a01a01a01ap-16);               // 1/8!
  p = fma(p, r, 0x1.a01a01a01a01ap-13);               // 1/7!
  p = fma(p, r, 0x1.6c16c16c16c17p-10);               // 1/6!
  p = fma(p, r, 0x1.1111111111111p-7);                // 1/5!
  p = fma(p, r, 0x1.5555555555555p-5);                // 1/4!
  p = fma(p, r, 0x1.5555555555555p-3);                // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const double ex = ldexp(p, (int)nf);
  double c, sn;
  cis(b, c, sn);
  ei = ex * sn;
  er = ex * c;
}

// near_faces (sysmat.f90:233-260) without FMA contraction, so the
// classification of borderline pairs follows the reference's arithmetic.
__device__ bool near_pair(const double *cpm, const double *cpn, double thr2) {
#pragma clang fp contract(off)
  double d0 = cpm[0] - cpn[0], d1 = cpm[1] - cpn[1], d2 = cpm[2] - cpn[2];
  double ds = d0 * d0 + d1 * d1 + d2 * d2;
  return ds < thr2;
}

// Adds one edge pair's D+H (ar, ai) and K (kr, ki) to every representation
// matrix: e (J,J), e detj Zsi (M,M), e detj K (M,J), -e K (J,M), the
// blocks of sysmat.f90:66-98 with the element's field action e, at the
// rows and columns the representation keeps.
// Each caller owns the entries it adds to (no other thread writes them
// during the launch), so the adds are plain read-modify-writes.
__device__ __forceinline__ void store_reps(const FillTarget &o, int64_t gi, int64_t gj, double ar, double ai,
                                           double kr, double ki) {
  const int64_t N = o.Noff;
  const double mr = o.zsi.x * ar - o.zsi.y * ai, mi = o.zsi.x * ai + o.zsi.y * ar;
  for (int r = 0; r < o.nrep; r++) {
    const int *mp = o.rep.map[r];
    const int a = mp[gi], b = mp[gj], am = mp[gi + N], bm = mp[gj + N];
    const double2 e = o.rep.e[r], ed = make_double2(e.x * o.detj, e.y * o.detj);
    const int64_t ld = o.rep.ld[r];
    double2 *Z = o.rep.Z[r];
    // adds w (vr + i vi); a real field action (mirror groups) keeps the
    // products of the real case
    auto add = [&](int i, int j, double2 w, double vr, double vi) {
      if (i < 0 || j < 0) return;
      double2 *z = Z + i + (int64_t)j * ld;
      const double2 v = *z;
      if (w.y == 0.0)
        *z = make_double2(v.x + w.x * vr, v.y + w.x * vi);
      else
        *z = make_double2(v.x + (w.x * vr - w.y * vi), v.y + (w.x * vi + w.y * vr));
    };
    add(a, b, e, ar, ai);
    add(am, bm, ed, mr, mi);
    add(am, b, ed, kr, ki);
    add(a, bm, e, -kr, -ki);
  }
}

// quad_tri_points (quad.f90:352-367): r = node(t,3) p1 + node(t,1) p2 + node(t,2) p3
__global__ void quad_points_kernel(int F, int Q, const double *__restrict__ p,
                                   const int *__restrict__ fnodes, QuadRule rule,
                                   double *__restrict__ qp) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= F * Q) return;
  int f = idx / Q, t = idx - f * Q;
  const double *p1 = p + 3 * fnodes[3 * f], *p2 = p + 3 * fnodes[3 * f + 1],
               *p3 = p + 3 * fnodes[3 * f + 2];
  double a = rule.nodes[2 * Q + t], b = rule.nodes[t], c = rule.nodes[Q + t];
  for (int k = 0; k < 3; k++) {
#pragma clang fp contract(off)
    qp[3 * idx + k] = a * p1[k] + b * p2[k] + c * p3[k];
  }
}

constexpr int kWaves = 4;

// Far / smooth kernel: one thread per (test face, source face) pair.
// QT > 0 fixes the rule size at compile time (full unrolling of the source
// point loop, so the exp/sincos chains of different points interleave);
// QT == 0 takes a.Q at run time.
// DK: the k-derivatives of the kernels instead (derivative of the fill with
// respect to the wavenumber): dG/dk = i R G, dg/dk = k G; smoothed parts
// dGs/dk = (i e^{ikR} + k R)/4pi, dgs/dk = k (e^{ikR} - 1)/(4pi R).
// LL: a lossless medium (Im k = 0), where e^{ikR} is cos kR + i sin kR with
// the same bits as cis_exp gives it.
template <bool SMOOTH, int QT, bool DK = false, bool LL = false>
__device__ __forceinline__ void pair_point_loop(const FarArgs &a, const double *__restrict__ tpts,
                                                const double *__restrict__ spts, int lane,
                                                double dx, double dy, double dz, double acc[36]) {
  // acc layout (complex pairs): 0 S0, 1-3 Srho, 4-6 Srho', 7 Srr',
  //                             8 M0, 9-11 Mrho, 12-14 Mrho', 15-17 Mx
  const int Q = QT > 0 ? QT : a.Q;
  const double kr = a.kr, ki = a.ki;
  const double k2r = kr * kr - ki * ki, k2i = 2.0 * kr * ki;
  for (int i = 0; i < Q; i++) {
    const double rx = tpts[3 * i], ry = tpts[3 * i + 1], rz = tpts[3 * i + 2];
    const double bx = dx + rx, by = dy + ry, bz = dz + rz;
    double T0r = 0, T0i = 0, T1r[3] = {0, 0, 0}, T1i[3] = {0, 0, 0};
    double U0r = 0, U0i = 0, U1r[3] = {0, 0, 0}, U1i[3] = {0, 0, 0};
#pragma unroll 2
    for (int j = 0; j < (QT > 0 ? QT : kMaxQ); j++) {
      if (QT == 0 && j >= Q) break;
      const double sx = spts[(3 * j) * 64 + lane], sy = spts[(3 * j + 1) * 64 + lane],
                   sz = spts[(3 * j + 2) * 64 + lane];
      const double x = bx - sx, y = by - sy, z = bz - sz;
      const double R2 = x * x + y * y + z * z;
      double Gr, Gi, gr, gi;
      if (!SMOOTH) {
        const double iR = rsqrt(R2);
        const double R = R2 * iR;
        double Er, Ei;
        if constexpr (LL)
          cis(kr * R, Er, Ei);
        else
          cis_exp(-ki * R, kr * R, Er, Ei);
        if constexpr (DK) {
          // dG/dk = i e^{ikR}/4pi, dg/dk = k G
          const double e = iR * kOofpi;
          const double G0r = e * Er, G0i = e * Ei;
          Gr = -Ei * kOofpi;
          Gi = Er * kOofpi;
          gr = kr * G0r - ki * G0i;
          gi = kr * G0i + ki * G0r;
        } else {
          const double e = iR * kOofpi;
          Gr = e * Er;
          Gi = e * Ei;
          // g = G * ((1/R^2 + ki/R) - i kr/R)
          const double ar = iR * iR + ki * iR, ai = -kr * iR;
          gr = Gr * ar - Gi * ai;
          gi = Gr * ai + Gi * ar;
        }
      } else if (R2 == 0.0) {
        // green.f90:33-34 (Gs = ik/4pi) and 88-91 (grad Gs = 0)
        Gr = DK ? 0.0 : -ki * kOofpi;
        Gi = DK ? kOofpi : kr * kOofpi;
        gr = 0.0;
        gi = 0.0;
      } else if (DK) {
        const double iR = rsqrt(R2);
        const double R = R2 * iR;
        double Er, Ei;
        if constexpr (LL)
          cis(kr * R, Er, Ei);
        else
          cis_exp(-ki * R, kr * R, Er, Ei);
        Gr = (-Ei + kr * R) * kOofpi;
        Gi = (Er + ki * R) * kOofpi;
        const double f = iR * kOofpi;
        gr = (kr * (Er - 1.0) - ki * Ei) * f;
        gi = (kr * Ei + ki * (Er - 1.0)) * f;
      } else {
        const double iR = rsqrt(R2);
        const double R = R2 * iR;
        double Er, Ei;
        if constexpr (LL)
          cis(kr * R, Er, Ei);
        else
          cis_exp(-ki * R, kr * R, Er, Ei);
        // Gs = [(e^{ikR} - 1)/R + k^2 R/2] / 4pi
        Gr = ((Er - 1.0) * iR + 0.5 * k2r * R) * kOofpi;
        Gi = (Ei * iR + 0.5 * k2i * R) * kOofpi;
        // grad Gs factor = [e^{ikR}(1/R^3 - ik/R^2) - 1/R^3 - k^2/(2R)] / 4pi
        const double iR2 = iR * iR, iR3 = iR2 * iR;
        const double ar = iR3 + ki * iR2, ai = -kr * iR2;
        gr = ((Er * ar - Ei * ai) - iR3 - 0.5 * k2r * iR) * kOofpi;
        gi = ((Er * ai + Ei * ar) - 0.5 * k2i * iR) * kOofpi;
      }
      const double wj = a.w[j];
      Gr *= wj;
      Gi *= wj;
      gr *= wj;
      gi *= wj;
      T0r += Gr;
      T0i += Gi;
      U0r += gr;
      U0i += gi;
      T1r[0] += Gr * sx; T1i[0] += Gi * sx;
      T1r[1] += Gr * sy; T1i[1] += Gi * sy;
      T1r[2] += Gr * sz; T1i[2] += Gi * sz;
      U1r[0] += gr * sx; U1i[0] += gi * sx;
      U1r[1] += gr * sy; U1i[1] += gi * sy;
      U1r[2] += gr * sz; U1i[2] += gi * sz;
    }
    const double wi = a.w[i];
    const double r3[3] = {rx, ry, rz};
    acc[0] += wi * T0r;
    acc[1] += wi * T0i;
    acc[16] += wi * U0r;
    acc[17] += wi * U0i;
    double dr = 0, di = 0;
    for (int c = 0; c < 3; c++) {
      acc[2 + 2 * c] += wi * r3[c] * T0r;
      acc[3 + 2 * c] += wi * r3[c] * T0i;
      acc[8 + 2 * c] += wi * T1r[c];
      acc[9 + 2 * c] += wi * T1i[c];
      dr += r3[c] * T1r[c];
      di += r3[c] * T1i[c];
      acc[18 + 2 * c] += wi * r3[c] * U0r;
      acc[19 + 2 * c] += wi * r3[c] * U0i;
      acc[24 + 2 * c] += wi * U1r[c];
      acc[25 + 2 * c] += wi * U1i[c];
    }
    acc[14] += wi * dr;
    acc[15] += wi * di;
    // rho_i x U1
    const double xr[3] = {ry * U1r[2] - rz * U1r[1], rz * U1r[0] - rx * U1r[2], rx * U1r[1] - ry * U1r[0]};
    const double xi[3] = {ry * U1i[2] - rz * U1i[1], rz * U1i[0] - rx * U1i[2], rx * U1i[1] - ry * U1i[0]};
    for (int c = 0; c < 3; c++) {
      acc[30 + 2 * c] += wi * xr[c];
      acc[31 + 2 * c] += wi * xi[c];
    }
  }
}

// Pair kernel: one thread per (test face m, source face n) with the test
// faces on the lanes (64 per wave) and four source faces per wave handled
// in turn (wave-uniform).  The 3x3 edge blocks of D+H (or the single
// operator asked for) and of K are written with plain, coalesced stores to
// face-side planes B[type][3p+q][n - n0][m] of the current slab of source
// faces; assemble_kernel then gathers them into the matrix.
struct PlaneArgs {
  FarArgs f;
  double2 *B;     // [2][9][S][F - mlo]
  int n0, S;      // slab of source faces [n0, n0 + S)
  int mlo;        // first test face: the pairs with both faces below n0 were done by earlier slabs
  int mode;       // 0: D+H and K; 1: D; 2: H; 3: K
};

// SPW source faces per wave (wave-uniform); the fill's autotuner picks SPW
// among 2, 4 and 8 per context.  The per-pair arithmetic does not depend on
// SPW, so every choice gives the same bits.
template <int QT, bool DK = false, int SPW = 4, bool LL = false>
__global__ __launch_bounds__(256, 2) void pair_plane_kernel(PlaneArgs pa) {
  constexpr int kSrcPerWave = SPW, kSrcPerBlock = kWaves * SPW;
  const FarArgs &a = pa.f;
  __shared__ double tpts[3 * kMaxQ * 64];          // test points, [3j+c][lane]
  __shared__ double spts[kSrcPerBlock][3 * kMaxQ];  // source points, broadcast
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int Q = QT > 0 ? QT : a.Q, F = a.F;
  const int m0 = pa.mlo + blockIdx.x * 64;
  const int W = F - pa.mlo;
  const int nb = pa.n0 + blockIdx.y * kSrcPerBlock;
  for (int idx = threadIdx.x; idx < 64 * Q * 3; idx += blockDim.x) {
    const int l = idx % 64, jc = idx / 64;
    const int m = m0 + l;
    double v = 0.0;
    if (m < F) {
      if (a.gen) {
        const int c = jc % 3;
        const double *q = a.qp + (int64_t)m * 3 * Q + (jc - c);
        const double *cm = a.cp + 3 * m;
        v = a.M[3 * c] * (q[0] - cm[0]) + a.M[3 * c + 1] * (q[1] - cm[1]) + a.M[3 * c + 2] * (q[2] - cm[2]);
      } else {
        v = a.js[jc % 3] * (a.qp[(int64_t)m * 3 * Q + jc] - a.cp[3 * m + jc % 3]);
      }
    }
    tpts[jc * 64 + l] = v;
  }
  for (int idx = threadIdx.x; idx < kSrcPerBlock * Q * 3; idx += blockDim.x) {
    const int t = idx / (3 * Q), jc = idx - t * 3 * Q;
    const int n = nb + t;
    double v = 0.0;
    if (n < pa.n0 + pa.S && n < F) v = a.qp[(int64_t)n * 3 * Q + jc] - a.cp[3 * n + jc % 3];
    spts[t][jc] = v;
  }
  __syncthreads();

  const int m = m0 + lane;
  const bool valid = m < F;
  const int mm = valid ? m : 0;
  double cmx, cmy, cmz;  // the (moved) test centroid
  if (a.gen) {
    const double *c = a.cp + 3 * mm;
    cmx = a.M[0] * c[0] + a.M[1] * c[1] + a.M[2] * c[2];
    cmy = a.M[3] * c[0] + a.M[4] * c[1] + a.M[5] * c[2];
    cmz = a.M[6] * c[0] + a.M[7] * c[1] + a.M[8] * c[2];
  } else {
    cmx = a.js[0] * a.cp[3 * mm];
    cmy = a.js[1] * a.cp[3 * mm + 1];
    cmz = a.js[2] * a.cp[3 * mm + 2];
  }
  const int64_t plane = (int64_t)pa.S * W;

  for (int t = 0; t < kSrcPerWave; t++) {
    const int tl = wave * kSrcPerWave + t;
    const int n = nb + tl;
    if (n >= pa.n0 + pa.S || n >= F) break;
    const double *cpn = a.cp + 3 * n;
    // Delta = c_m - c_n.  The point loop runs over the source points in its
    // outer ("i") index and the test points on the lanes ("j"), so it is
    // given -Delta; its moments then come out with the roles exchanged:
    // acc[2..7] = S_rho' (source), acc[8..13] = S_rho (test),
    // acc[18..23] = M_rho', acc[24..29] = M_rho, acc[30..35] = -M_x.
    const double dx = cmx - cpn[0], dy = cmy - cpn[1], dz = cmz - cpn[2];
    bool near;
    if (a.gen) {
      // the reference's |c_m - J c_n| with the host list's rounding
      double d[3];
      image_near_diff(a.J, a.ccls + 3 * mm, a.ccls + 3 * n, d);
      const double zero[3] = {0.0, 0.0, 0.0};
      near = near_pair(d, zero, a.thr2);
    } else {
      const double clm[3] = {a.js[0] * a.ccls[3 * mm], a.js[1] * a.ccls[3 * mm + 1], a.js[2] * a.ccls[3 * mm + 2]};
      near = near_pair(clm, a.ccls + 3 * n, a.thr2);
    }
    double acc[36];
    for (int k = 0; k < 36; k++) acc[k] = 0.0;
    if (near)
      pair_point_loop<true, QT, DK, LL>(a, spts[tl], tpts, lane, -dx, -dy, -dz, acc);
    else
      pair_point_loop<false, QT, DK, LL>(a, spts[tl], tpts, lane, -dx, -dy, -dz, acc);
    if (!valid) continue;

    const double AA = a.area[mm] * a.area[n];
    const bool self = (mm == n) && !a.image;
    const double D3[3] = {dx, dy, dz};
    const int nl = n - pa.n0;
#pragma unroll 1
    for (int q = 0; q < 3; q++) {
      const int sq = 3 * n + q;
      const double coefn = a.side.coef[sq], divn = a.side.div[sq];
      const double vn[3] = {a.side.v[3 * sq] - cpn[0], a.side.v[3 * sq + 1] - cpn[1], a.side.v[3 * sq + 2] - cpn[2]};
#pragma unroll 1
      for (int p = 0; p < 3; p++) {
        const int sp = 3 * mm + p;
        const double coefm = a.side.coef[sp], divm = a.side.div[sp];
        double vm[3];
        if (a.gen) {
          const double *v = a.side.v + 3 * sp;
          vm[0] = (a.M[0] * v[0] + a.M[1] * v[1] + a.M[2] * v[2]) - cmx;
          vm[1] = (a.M[3] * v[0] + a.M[4] * v[1] + a.M[5] * v[2]) - cmy;
          vm[2] = (a.M[6] * v[0] + a.M[7] * v[1] + a.M[8] * v[2]) - cmz;
        } else {
          vm[0] = a.js[0] * a.side.v[3 * sp] - cmx;
          vm[1] = a.js[1] * a.side.v[3 * sp + 1] - cmy;
          vm[2] = a.js[2] * a.side.v[3 * sp + 2] - cmz;
        }
        // D: Srr' - nu_n.Srho - nu_m.Srho' + (nu_m.nu_n) S0
        const double vmn = vm[0] * vn[0] + vm[1] * vn[1] + vm[2] * vn[2];
        double sr = acc[14] + vmn * acc[0], si = acc[15] + vmn * acc[1];
        for (int c = 0; c < 3; c++) {
          sr -= vn[c] * acc[8 + 2 * c] + vm[c] * acc[2 + 2 * c];
          si -= vn[c] * acc[9 + 2 * c] + vm[c] * acc[3 + 2 * c];
        }
        const double fD = AA * coefm * coefn;
        const double Dr = fD * (a.c1.x * sr - a.c1.y * si);
        const double Di = fD * (a.c1.x * si + a.c1.y * sr);
        const double fH = AA * divm * divn;
        const double Hr = fH * (a.c2.x * acc[0] - a.c2.y * acc[1]);
        const double Hi = fH * (a.c2.x * acc[1] + a.c2.y * acc[0]);
        double Kr = 0, Ki = 0;
        if (!self && (pa.mode == 0 || pa.mode == 3)) {
          // det[rho - nu_m, Delta + rho - rho', rho' - nu_n] summed:
          //  -(Delta + nu_m - nu_n).Mx - [(Delta x nu_n) - (nu_n x nu_m)].Mrho
          //  - [(nu_m x Delta) + (nu_n x nu_m)].Mrho' + nu_m.(Delta x nu_n) M0
          double av[3], dn[3], nm[3], md[3];
          for (int c = 0; c < 3; c++) av[c] = D3[c] + vm[c] - vn[c];
          dn[0] = D3[1] * vn[2] - D3[2] * vn[1];
          dn[1] = D3[2] * vn[0] - D3[0] * vn[2];
          dn[2] = D3[0] * vn[1] - D3[1] * vn[0];
          nm[0] = vn[1] * vm[2] - vn[2] * vm[1];
          nm[1] = vn[2] * vm[0] - vn[0] * vm[2];
          nm[2] = vn[0] * vm[1] - vn[1] * vm[0];
          md[0] = vm[1] * D3[2] - vm[2] * D3[1];
          md[1] = vm[2] * D3[0] - vm[0] * D3[2];
          md[2] = vm[0] * D3[1] - vm[1] * D3[0];
          const double c0 = vm[0] * dn[0] + vm[1] * dn[1] + vm[2] * dn[2];
          double kr = c0 * acc[16], ki = c0 * acc[17];
          for (int c = 0; c < 3; c++) {
            const double cr = dn[c] - nm[c], cs = md[c] + nm[c];
            kr += av[c] * acc[30 + 2 * c] - cr * acc[24 + 2 * c] - cs * acc[18 + 2 * c];
            ki += av[c] * acc[31 + 2 * c] - cr * acc[25 + 2 * c] - cs * acc[19 + 2 * c];
          }
          Kr = fD * kr;
          Ki = fD * ki;
        }
        double vr, vi;
        if (pa.mode == 1) {
          vr = Dr;
          vi = Di;
        } else if (pa.mode == 2) {
          vr = Hr;
          vi = Hi;
        } else {
          vr = Dr + Hr;
          vi = Di + Hi;
        }
        const int64_t off = (int64_t)(3 * p + q) * plane + (int64_t)nl * W + (mm - pa.mlo);
        pa.B[off] = make_double2(vr, vi);
        pa.B[9 * plane + off] = make_double2(Kr, Ki);
      }
    }
  }
}

// Gathers the face-side planes of one slab into the matrix.  A slab holds
// the blocks of the face pairs (test m >= n0, source n in [n0, n0 + S)).
// The far and smoothed-kernel integrals are symmetric in the two faces
// (G(r, r') = G(r', r); the D, H and K bilinear forms are symmetric under
// exchanging the faces, the sign of grad' G cancelling against the order of
// the triple product), so the block of (test n, source m) is the transpose
// of that of (test m, source n), and every unordered pair is integrated
// once.  Direct launch: thread (row r of the output numbering, source edge
// e2 of the slab) sums the contributions of the (<= 2) test face sides of
// r's edge with a face >= n0 and the (<= 2) source face sides of e2 that
// lie in the slab.  Transposed launch: thread (test edge e1 of the slab,
// source edge e2 with a face beyond the slab) adds the transposed blocks of
// the test sides of e1 in the slab and the source sides of e2 beyond it.
// Either updates its entry with one read-modify-write (each entry is owned
// by one thread per launch).
// Adds the D+H (ar, ai) and K (kr, ki) sums of edge pair (gi, gj) to the
// output: the four PMCHWT blocks (mode 0) or the single operator.
__device__ __forceinline__ void store_entry(const FillTarget &o, int64_t gi, int64_t gj, double ar, double ai,
                                                 double kr, double ki) {
  if (o.nrep > 0) {
    store_reps(o, gi, gj, ar, ai, kr, ki);
  } else if (o.mode == 0) {
    const int64_t N = o.Noff, ld = o.ld;
    double2 *z = o.Z + gi + gj * ld;
    double2 v = *z;
    *z = make_double2(v.x + ar, v.y + ai);
    z = o.Z + (gi + N) + (gj + N) * ld;
    v = *z;
    *z = make_double2(v.x + (o.zsi.x * ar - o.zsi.y * ai), v.y + (o.zsi.x * ai + o.zsi.y * ar));
    z = o.Z + (gi + N) + gj * ld;
    v = *z;
    *z = make_double2(v.x + kr, v.y + ki);
    z = o.Z + gi + (gj + N) * ld;
    v = *z;
    *z = make_double2(v.x - kr, v.y - ki);
  } else {
    double2 *z = o.Z + gi + gj * o.ld;
    double2 v = *z;
    if (o.mode == 3)
      *z = make_double2(v.x + kr, v.y + ki);
    else
      *z = make_double2(v.x + ar, v.y + ai);
  }
}

struct AssembleArgs {
  const double2 *B;
  int F, S, n0, mlo;
  const int *row_local;   // output row -> local edge of this mesh, -1 if none
  int64_t nrows;
  const int *slab_edges;  // local source edges of this slab
  const int *eside;       // [E][2] face-side codes 3f+s of each local edge, -1 if none
  const int *col_index;   // local edge -> output column
  double ktr;             // factor of K in the transposed blocks: det J of the image (1 without)
  FillTarget out;
};

__global__ __launch_bounds__(256) void assemble_kernel(AssembleArgs g) {
  const int64_t r = blockIdx.x * 256 + (int64_t)threadIdx.x;
  if (r >= g.nrows) return;
  const int e1 = g.row_local[r];
  if (e1 < 0) return;
  const int e2 = g.slab_edges[blockIdx.y];
  const int W = g.F - g.mlo;
  const int64_t plane = (int64_t)g.S * W;
  const int s1[2] = {g.eside[2 * e1], g.eside[2 * e1 + 1]};
  const int s2[2] = {g.eside[2 * e2], g.eside[2 * e2 + 1]};
  double ar = 0, ai = 0, kr = 0, ki = 0;
  bool any = false;
#pragma unroll
  for (int b = 0; b < 2; b++) {
    if (s2[b] < 0) continue;
    const int n = s2[b] / 3, q = s2[b] - 3 * n;
    const int nl = n - g.n0;
    if (nl < 0 || nl >= g.S) continue;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      if (s1[a] < 0) continue;
      const int m = s1[a] / 3, p = s1[a] - 3 * m;
      if (m < g.mlo) continue;  // integrated by the slab of m, added there transposed
      const int64_t off = (int64_t)(3 * p + q) * plane + (int64_t)nl * W + (m - g.mlo);
      const double2 v = g.B[off], k = g.B[9 * plane + off];
      ar += v.x;
      ai += v.y;
      kr += k.x;
      ki += k.y;
      any = true;
    }
  }
  // rows whose test faces all precede the slab (about half of them, on
  // average over the slabs) receive nothing here: no read-modify-write
  if (any) store_entry(g.out, r, g.col_index[e2], ar, ai, kr, ki);
}

// Transposed launch, tiled: a workgroup takes kTR test edges of the slab x
// kTC source edges beyond it.  The plane reads run along the source faces
// (lanes over source edges), the matrix updates along the rows (lanes over
// test edges); the sums pass through LDS between the two phases.
constexpr int kTR = 32, kTC = 64;

__global__ __launch_bounds__(256) void assemble_transposed_kernel(AssembleArgs g, const int *__restrict__ cols,
                                                                  int ne, int nc) {
  __shared__ double2 sa[kTR][kTC + 1], sk[kTR][kTC + 1];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * kTR, c0 = blockIdx.y * kTC;
  const int W = g.F - g.mlo;
  const int64_t plane = (int64_t)g.S * W;
  {
    const int c = tid % kTC, rsub = tid / kTC;  // 4 rows per pass
    const int cc = c0 + c;
    const int e2 = cc < nc ? cols[cc] : -1;
    int s2[2] = {-1, -1};
    if (e2 >= 0) {
      s2[0] = g.eside[2 * e2];
      s2[1] = g.eside[2 * e2 + 1];
    }
    for (int rr = rsub; rr < kTR; rr += 256 / kTC) {
      const int t = r0 + rr;
      double ar = 0, ai = 0, kr = 0, ki = 0;
      if (t < ne && e2 >= 0) {
        const int e1 = g.slab_edges[t];  // test edge with a side in the slab
        const int s1[2] = {g.eside[2 * e1], g.eside[2 * e1 + 1]};
#pragma unroll
        for (int a = 0; a < 2; a++) {
          if (s1[a] < 0) continue;
          const int fa = s1[a] / 3, pa = s1[a] - 3 * fa;
          const int al = fa - g.n0;
          if (al < 0 || al >= g.S) continue;
#pragma unroll
          for (int b = 0; b < 2; b++) {
            if (s2[b] < 0) continue;
            const int fb = s2[b] / 3, qb = s2[b] - 3 * fb;
            if (fb < g.n0 + g.S) continue;
            // block of (test fb, source fa), entry (side qb, side pa)
            const int64_t off = (int64_t)(3 * qb + pa) * plane + (int64_t)al * W + (fb - g.mlo);
            const double2 v = g.B[off], k = g.B[9 * plane + off];
            ar += v.x;
            ai += v.y;
            kr += k.x;
            ki += k.y;
          }
        }
      }
      sa[rr][c] = make_double2(ar, ai);
      sk[rr][c] = make_double2(kr, ki);
    }
  }
  __syncthreads();
  {
    const int rr = tid % kTR, csub = tid / kTR;  // 8 columns per pass
    const int t = r0 + rr;
    if (t < ne) {
      const int64_t gi = g.col_index[g.slab_edges[t]];
      for (int c = csub; c < kTC; c += 256 / kTR) {
        const int cc = c0 + c;
        if (cc >= nc) break;
        const double2 v = sa[rr][c], k = sk[rr][c];
        store_entry(g.out, gi, g.col_index[cols[cc]], v.x, v.y, g.ktr * k.x, g.ktr * k.y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Analytic singular integrals (int.f90, Hanninen 2006), FP64, evaluated with
// FMA contraction off so that exact-zero branch tests and rounding follow the
// reference's arithmetic.  They are templates on the scalar type: double for
// the fill, D1 (a value with its derivative in one coordinate, forward mode)
// for the exact derivatives of the node-coordinate adjoint.  Branches test
// the value (dval), so both instantiations take the same branch.
struct D1 {
  double v, d;
  D1() = default;
  __device__ D1(double x) : v(x), d(0.0) {}
  __device__ D1(double x, double y) : v(x), d(y) {}
};
__device__ inline D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
__device__ inline D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
__device__ inline D1 operator-(D1 a) { return D1(-a.v, -a.d); }
__device__ inline D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline D1 operator*(double a, D1 b) { return D1(a * b.v, a * b.d); }
__device__ inline D1 operator*(D1 a, double b) { return D1(a.v * b, a.d * b); }
__device__ inline D1 operator/(D1 a, D1 b) {
  const double q = a.v / b.v;
  return D1(q, (a.d - q * b.d) / b.v);
}
__device__ inline D1 operator/(D1 a, double b) { return D1(a.v / b, a.d / b); }
__device__ inline D1 operator/(double a, D1 b) {
  const double q = a / b.v;
  return D1(q, -q * b.d / b.v);
}
__device__ inline D1 operator+(D1 a, double b) { return D1(a.v + b, a.d); }
__device__ inline D1 operator+(double a, D1 b) { return D1(a + b.v, b.d); }
__device__ inline D1 operator-(D1 a, double b) { return D1(a.v - b, a.d); }
__device__ inline D1 operator-(double a, D1 b) { return D1(a - b.v, -b.d); }
__device__ inline D1 &operator+=(D1 &a, D1 b) { return a = a + b; }
__device__ inline double dval(double x) { return x; }
__device__ inline double dval(D1 x) { return x.v; }
__device__ inline double tsqrt(double x) { return sqrt(x); }
__device__ inline D1 tsqrt(D1 x) {
  const double r = sqrt(x.v);
  return D1(r, 0.5 * x.d / r);
}
__device__ inline double tlog(double x) { return log(x); }
__device__ inline D1 tlog(D1 x) { return D1(log(x.v), x.d / x.v); }
__device__ inline double tfabs(double x) { return fabs(x); }
__device__ inline D1 tfabs(D1 x) { return x.v < 0.0 ? -x : x; }
__device__ inline double tatan2(double y, double x) { return atan2(y, x); }
__device__ inline D1 tatan2(D1 y, D1 x) {
  return D1(atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / (x.v * x.v + y.v * y.v));
}

template <class T>
struct NearFaceT {
  T p[3][3];  // nodes
  T n[3];     // normal
  T m[3][3];  // in-plane edge normals m(:,e)
  T s[3][3];  // edge tangents s(:,e)
};
using NearFace = NearFaceT<double>;

// the type of a value computed from a TA and a TB (D1 if either is)
template <class TA, class TB>
struct Promote {
  using type = D1;
};
template <>
struct Promote<double, double> {
  using type = double;
};
template <class TA, class TB>
using PromoteT = typename Promote<TA, TB>::type;

template <class TA, class TB>
__device__ inline PromoteT<TA, TB> dot3(const TA *a, const TB *b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class T>
__device__ inline T norm3(const T *a) {
#pragma clang fp contract(off)
  return tsqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}
template <class TA, class TB, class TR>
__device__ inline void cross3(const TA *a, const TB *b, TR *r) {
#pragma clang fp contract(off)
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// The closed forms take the point r (TP) and the face (TF) separately, so
// that a derivative in one of them leaves the other in plain doubles.

// int.f90:15-53 intLm1 for side e (nodes e, e+1)
template <class TP, class TF>
__device__ PromoteT<TP, TF> d_intLm1(const TP *r, const NearFaceT<TF> &f, int e) {
#pragma clang fp contract(off)
  using T = PromoteT<TP, TF>;
  const TF *p1 = f.p[e], *p2 = f.p[(e + 1) % 3], *s = f.s[e];
  T d2[3] = {r[0] - p2[0], r[1] - p2[1], r[2] - p2[2]};
  T d1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  T Rp = norm3(d2), Rm = norm3(d1);
  T q2[3] = {p2[0] - r[0], p2[1] - r[1], p2[2] - r[2]};
  T q1[3] = {p1[0] - r[0], p1[1] - r[1], p1[2] - r[2]};
  T sp = dot3(q2, s), sm = dot3(q1, s);
  T div1 = Rp - sp, div2 = Rm + sm;
  if (fabs(dval(div1)) > fabs(dval(div2))) return tlog((Rm - sm) / div1);
  return tlog((Rp + sp) / div2);
}

// int.f90:55-135 intL1 and intL3 of side e (they share R0^2, s+-, R+-)
template <class TP, class TF, class T>
__device__ void d_intL13(const TP *r, const NearFaceT<TF> &f, int e, T lm1, T *L1, T *L3) {
#pragma clang fp contract(off)
  const TF *p1 = f.p[e], *p2 = f.p[(e + 1) % 3], *s = f.s[e];
  TF p[3];
  T rp[3];
  for (int k = 0; k < 3; k++) p[k] = 0.5 * (p1[k] + p2[k]);
  for (int k = 0; k < 3; k++) rp[k] = r[k] - p[k];
  T h = dot3(f.n, rp), t = dot3(f.m[e], rp);
  T R02 = t * t + h * h;
  T d2[3] = {r[0] - p2[0], r[1] - p2[1], r[2] - p2[2]};
  T d1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  T Rp = norm3(d2), Rm = norm3(d1);
  T q2[3] = {p2[0] - r[0], p2[1] - r[1], p2[2] - r[2]};
  T q1[3] = {p1[0] - r[0], p1[1] - r[1], p1[2] - r[2]};
  T sp = dot3(q2, s), sm = dot3(q1, s);
  T l1;
  if (dval(R02) == 0)
    l1 = 0.5 * (sp * Rp - sm * Rm);
  else
    l1 = 0.5 * R02 * lm1 + 0.5 * (sp * Rp - sm * Rm);
  *L1 = l1;
  *L3 = 0.75 * R02 * l1 + 0.25 * (sp * (Rp * Rp * Rp) - sm * (Rm * Rm * Rm));
}

// int.f90:137-162 intSm3h (signed solid angle)
template <class TP, class TF>
__device__ PromoteT<TP, TF> d_intSm3h(const TP *r, const NearFaceT<TF> &f) {
#pragma clang fp contract(off)
  using T = PromoteT<TP, TF>;
  T a[3][3];
  for (int v = 0; v < 3; v++) {
    T d[3] = {f.p[v][0] - r[0], f.p[v][1] - r[1], f.p[v][2] - r[2]};
    T l = norm3(d);
    for (int k = 0; k < 3; k++) a[v][k] = d[k] / l;
  }
  T x = 1.0 + dot3(a[0], a[1]) + dot3(a[0], a[2]) + dot3(a[1], a[2]);
  T c[3];
  cross3(a[1], a[2], c);
  T y = tfabs(dot3(a[0], c));
  T om = 2.0 * tatan2(y, x);
  T rp1[3] = {r[0] - f.p[0][0], r[1] - f.p[0][1], r[2] - f.p[0][2]};
  return dval(dot3(f.n, rp1)) > 0.0 ? tfabs(om) : -tfabs(om);
}

template <class T>
struct PointIntsT {
  T h;          // n.(r - p1)
  T lm1[3], l1[3], l3[3];
  T sm3h, sm1, s1;
};
using PointInts = PointIntsT<double>;

template <class TP, class TF>
__device__ void d_point_ints(const TP *r, const NearFaceT<TF> &f, PointIntsT<PromoteT<TP, TF>> &o) {
#pragma clang fp contract(off)
  using T = PromoteT<TP, TF>;
  T t[3];
  for (int e = 0; e < 3; e++) {
    T d[3] = {r[0] - f.p[e][0], r[1] - f.p[e][1], r[2] - f.p[e][2]};
    t[e] = dot3(f.m[e], d);
  }
  T d0[3] = {r[0] - f.p[0][0], r[1] - f.p[0][1], r[2] - f.p[0][2]};
  const T h = dot3(f.n, d0);
  o.h = h;
  for (int e = 0; e < 3; e++) {
    o.lm1[e] = d_intLm1(r, f, e);
    d_intL13(r, f, e, o.lm1[e], &o.l1[e], &o.l3[e]);
  }
  o.sm3h = d_intSm3h(r, f);
  // int.f90:181-226 intSm1
  if (dval(h) == 0) {
    T e1 = dval(t[0]) == 0 ? T(0.0) : o.lm1[0];
    T e2 = dval(t[1]) == 0 ? T(0.0) : o.lm1[1];
    T e3 = dval(t[2]) == 0 ? T(0.0) : o.lm1[2];
    o.sm1 = -(t[0] * e1 + t[1] * e2 + t[2] * e3);
  } else {
    o.sm1 = -h * o.sm3h - (t[0] * o.lm1[0] + t[1] * o.lm1[1] + t[2] * o.lm1[2]);
  }
  // int.f90:228-246 intS1
  o.s1 = (h * h) / 3.0 * o.sm1 - 1.0 / 3.0 * (t[0] * o.l1[0] + t[1] * o.l1[1] + t[2] * o.l1[2]);
}

struct NearArgs {
  int Q;
  double w[kMaxQ];
  const int *pairs;  // (m, n)
  int64_t npairs;
  const double *qp, *p, *fn, *fm, *fs, *area;
  const int *fnodes;
  FaceSide side;
  int mode;     // output of FillTarget::mode
  double2 kk4;  // 1/(4pi) (real, kept complex for symmetry)
  double2 kk8;  // k^2/(8pi)
  double2 c1, c2;
  double js[3];  // test faces reflected by diag(js)
  int image;
  int gen;       // test points and free vertices moved by M = J^T (non-diagonal image)
  double M[9];
  double2 *nb;   // per pair and (test side, source side): D+H (or the selected operator), K
};

// M x for the near rule's test points (MATMUL(TRANSPOSE(J), r) in the
// reference's order, no contraction)
__device__ __forceinline__ void mat3_apply(const double *M, const double *x, double *y) {
#pragma clang fp contract(off)
  for (int i = 0; i < 3; i++) {
    double t = M[3 * i] * x[0];
    t = t + M[3 * i + 1] * x[1];
    t = t + M[3 * i + 2] * x[2];
    y[i] = t;
  }
}

__global__ __launch_bounds__(128) void near_fill_kernel(NearArgs a) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= a.npairs) return;
  const int m = a.pairs[2 * id], n = a.pairs[2 * id + 1];
  const bool self = (m == n) && !a.image;
  NearFace f;
  for (int v = 0; v < 3; v++) {
    int node = a.fnodes[3 * n + v];
    for (int k = 0; k < 3; k++) f.p[v][k] = a.p[3 * node + k];
  }
  for (int k = 0; k < 3; k++) f.n[k] = a.fn[3 * n + k];
  for (int e = 0; e < 3; e++)
    for (int k = 0; k < 3; k++) {
      f.m[e][k] = a.fm[9 * n + 3 * e + k];
      f.s[e][k] = a.fs[9 * n + 3 * e + k];
    }
  double coefn[3], divn[3], vn[3][3];
  for (int q = 0; q < 3; q++) {
    coefn[q] = a.side.coef[3 * n + q];
    divn[q] = a.side.div[3 * n + q];
    for (int k = 0; k < 3; k++) vn[q][k] = a.side.v[3 * (3 * n + q) + k];
  }
  double coefm[3], vm[3][3];
  for (int p = 0; p < 3; p++) {
    coefm[p] = a.side.coef[3 * m + p];
    if (a.gen)
      mat3_apply(a.M, a.side.v + 3 * (3 * m + p), vm[p]);
    else
      for (int k = 0; k < 3; k++) vm[p][k] = a.js[k] * a.side.v[3 * (3 * m + p) + k];
  }
  double accD[3][3][2], accK[3][3][2], accH[3][2];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) accD[p][q][0] = accD[p][q][1] = accK[p][q][0] = accK[p][q][1] = 0.0;
  for (int q = 0; q < 3; q++) accH[q][0] = accH[q][1] = 0.0;

  for (int i = 0; i < a.Q; i++) {
    const double *rq = a.qp + ((int64_t)m * a.Q + i) * 3;
    double r[3];
    if (a.gen) {
      mat3_apply(a.M, rq, r);
    } else {
      r[0] = a.js[0] * rq[0];
      r[1] = a.js[1] * rq[1];
      r[2] = a.js[2] * rq[2];
    }
    PointInts P;
    d_point_ints(r, f, P);
    // intK31 / intK3m1 (int.f90:330-374), shared by the three edges
    double k31[3], k3m1[3];
    for (int c = 0; c < 3; c++) {
      double s = f.m[0][c] * P.l1[0] + f.m[1][c] * P.l1[1] + f.m[2][c] * P.l1[2];
      k31[c] = (P.h == 0) ? s : s - P.h * f.n[c] * P.sm1;
      k3m1[c] = f.m[0][c] * P.lm1[0] + f.m[1][c] * P.lm1[1] + f.m[2][c] * P.lm1[2] + f.n[c] * P.sm3h;
    }
    double rho[3];
    for (int c = 0; c < 3; c++) rho[c] = r[c] - P.h * f.n[c];
    const double wi = a.w[i];
    double fm_[3][3];
    for (int p = 0; p < 3; p++)
      for (int c = 0; c < 3; c++) fm_[p][c] = coefm[p] * (r[c] - vm[p][c]);
    for (int q = 0; q < 3; q++) {
      // intK1m1 / intK11 (int.f90:248-274): sign L/A * {Sm1, S1}
      const double h1 = divn[q] * P.sm1, h0 = divn[q] * P.s1;
      const double aHr = a.kk4.x * h1 - (a.kk8.x * h0), aHi = -(a.kk8.y * h0);
      accH[q][0] += wi * aHr;
      accH[q][1] += wi * aHi;
      // intK2m1 / intK21 (int.f90:276-328)
      double k2m1[3], k21[3];
      for (int c = 0; c < 3; c++) {
        k2m1[c] = coefn[q] * (f.m[0][c] * P.l1[0] + f.m[1][c] * P.l1[1] + f.m[2][c] * P.l1[2] +
                              (rho[c] - vn[q][c]) * P.sm1);
        k21[c] = coefn[q] * (1.0 / 3.0 * (f.m[0][c] * P.l3[0] + f.m[1][c] * P.l3[1] + f.m[2][c] * P.l3[2]) +
                             (rho[c] - vn[q][c]) * P.s1);
      }
      double aD[3][2];
      for (int c = 0; c < 3; c++) {
        aD[c][0] = a.kk4.x * k2m1[c] - a.kk8.x * k21[c];
        aD[c][1] = -(a.kk8.y * k21[c]);
      }
      double aK[3][2] = {{0, 0}, {0, 0}, {0, 0}};
      if (!self) {
        // intK4m1 / intK41 (int.f90:376-406): -sign L/(2A) (r - p) x {K3m1, K31}
        double d[3] = {r[0] - vn[q][0], r[1] - vn[q][1], r[2] - vn[q][2]};
        double c1v[3], c0v[3];
        cross3(d, k3m1, c1v);
        cross3(d, k31, c0v);
        const double sc = -coefn[q];
        for (int c = 0; c < 3; c++) {
          const double k4m1 = sc * c1v[c], k41 = sc * c0v[c];
          aK[c][0] = a.kk4.x * k4m1 - a.kk8.x * k41;
          aK[c][1] = -(a.kk8.y * k41);
        }
      }
      for (int p = 0; p < 3; p++) {
        double dr = 0, di = 0, kr = 0, ki = 0;
        for (int c = 0; c < 3; c++) {
          dr += fm_[p][c] * aD[c][0];
          di += fm_[p][c] * aD[c][1];
          kr += fm_[p][c] * aK[c][0];
          ki += fm_[p][c] * aK[c][1];
        }
        accD[p][q][0] += wi * dr;
        accD[p][q][1] += wi * di;
        accK[p][q][0] += wi * kr;
        accK[p][q][1] += wi * ki;
      }
    }
  }
  const double Am = a.area[m];
  double2 *out = a.nb + 18 * id;
  for (int p = 0; p < 3; p++) {
    const double divm = a.side.div[3 * m + p];
    for (int q = 0; q < 3; q++) {
      const double dr = accD[p][q][0] * Am, di = accD[p][q][1] * Am;
      const double Dr = a.c1.x * dr - a.c1.y * di, Di = a.c1.x * di + a.c1.y * dr;
      const double hr = accH[q][0] * Am * divm, hi = accH[q][1] * Am * divm;
      const double Hr = a.c2.x * hr - a.c2.y * hi, Hi = a.c2.x * hi + a.c2.y * hr;
      const double Kr = accK[p][q][0] * Am, Ki = accK[p][q][1] * Am;
      double2 v0;
      if (a.mode == 1)
        v0 = make_double2(Dr, Di);
      else if (a.mode == 2)
        v0 = make_double2(Hr, Hi);
      else if (a.mode == 3)
        v0 = make_double2(Kr, Ki);
      else
        v0 = make_double2(Dr + Hr, Di + Hi);
      out[2 * (3 * p + q)] = v0;
      out[2 * (3 * p + q) + 1] = make_double2(Kr, Ki);
    }
  }
}

// Index of the near pair (m, n) in the list sorted by test then source face
// (off[m] .. off[m+1] the pairs of test face m), -1 if (m, n) is not near.
__device__ __forceinline__ int64_t find_near(const int *__restrict__ pairs, const int64_t *__restrict__ off, int m,
                                             int n) {
  int64_t lo = off[m], hi = off[m + 1];
  const int64_t end = hi;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (pairs[2 * mid + 1] < n)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && pairs[2 * lo + 1] == n) ? lo : -1;
}

struct NearEmitArgs {
  const int *pairs;
  const int64_t *off;
  int64_t npairs;
  const int *eside;  // face-side codes 3 f + s of each local edge's two faces (-1 none)
  const int *s_l;    // local edge of each face side
  FaceSide side;     // output index of each face side
  const double2 *nb;
  int mode;
  FillTarget out;
};

// The near pairs' analytic parts added to the matrix in a fixed order,
// without atomics.  An entry (edge i, edge j) collects the pairs (m, n) with
// m a face of i and n a face of j; thread (pair, test side, source side)
// owns the entry when its pair is the first near one in the order (face of
// i, face of j) = (+,+), (+,-), (-,+), (-,-), and then sums all of them in
// that order and adds the sum once.
__global__ __launch_bounds__(256) void near_emit_kernel(NearEmitArgs a) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 9 * a.npairs) return;
  const int64_t id = t / 9;
  const int pq = (int)(t - 9 * id), p = pq / 3, q = pq % 3;
  const int m = a.pairs[2 * id], n = a.pairs[2 * id + 1];
  const int sm = 3 * m + p, sn = 3 * n + q;
  const int i = a.s_l[sm], j = a.s_l[sn];
  double2 v0 = make_double2(0.0, 0.0), v1 = make_double2(0.0, 0.0);
  bool first = true;
  for (int ci = 0; ci < 2; ci++) {
    const int ca = a.eside[2 * i + ci];
    if (ca < 0) continue;
    for (int cj = 0; cj < 2; cj++) {
      const int cb = a.eside[2 * j + cj];
      if (cb < 0) continue;
      const int64_t k = find_near(a.pairs, a.off, ca / 3, cb / 3);
      if (k < 0) continue;
      if (first && k != id) return;  // another thread owns this entry
      first = false;
      const int slot = 3 * (ca % 3) + cb % 3;
      const double2 u0 = a.nb[18 * k + 2 * slot], u1 = a.nb[18 * k + 2 * slot + 1];
      v0.x += u0.x;
      v0.y += u0.y;
      v1.x += u1.x;
      v1.y += u1.y;
    }
  }
  const int gi = a.side.g[sm], gj = a.side.g[sn];
  const FillTarget &o = a.out;
  if (o.nrep > 0) {
    store_reps(o, gi, gj, v0.x, v0.y, v1.x, v1.y);
    return;
  }
  auto add = [&](int64_t r, int64_t c, double vr, double vi) {
    double2 *z = o.Z + r + c * o.ld;
    const double2 w = *z;
    *z = make_double2(w.x + vr, w.y + vi);
  };
  if (a.mode == 0) {
    const int64_t N = o.Noff;
    add(gi, gj, v0.x, v0.y);
    add(gi + N, gj + N, o.zsi.x * v0.x - o.zsi.y * v0.y, o.zsi.x * v0.y + o.zsi.y * v0.x);
    add(gi + N, gj, v1.x, v1.y);
    add(gi, gj + N, -v1.x, -v1.y);
  } else {
    add(gi, gj, v0.x, v0.y);
  }
}

static int quad_points_into(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double *qp) {
  int total = (int)(m->F * rule.Q);
  hipLaunchKernelGGL(quad_points_kernel, dim3((total + 255) / 256), dim3(256), 0, ctx->stream,
                     (int)m->F, rule.Q, m->p.as<double>(), m->fnodes.as<int>(), rule, qp);
  MOM_HIP(hipGetLastError());
  return 0;
}

int compute_quad_points(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule) {
  size_t need = (size_t)m->F * rule.Q * 3 * sizeof(double);
  if (ctx->qp.bytes < need) {
    int rc = ctx->qp.alloc(need);
    if (rc) return rc;
  }
  return quad_points_into(ctx, m, rule, ctx->qp.as<double>());
}

static FaceSide sides(const mom_mesh *m, bool global) {
  FaceSide s;
  s.g = global ? m->s_g.as<int>() : m->s_l.as<int>();
  s.coef = m->s_coef.as<double>();
  s.div = m->s_div.as<double>();
  s.v = m->s_v.as<double>();
  return s;
}

// Event pairs around the fill's launches while ctx->fill_profile is on;
// harvest() (after the stream has drained) adds their times to
// ctx->fill_prof.  Kind 0: pair kernel, 1: assembly, 2: near part.
struct FillTimer {
  mom_ctx *ctx;
  struct Rec {
    hipEvent_t e0, e1;
    int kind;
    double points;
  };
  std::vector<Rec> recs;
  explicit FillTimer(mom_ctx *c) : ctx(c) {}
  hipEvent_t begin() {
    if (!ctx->fill_profile) return nullptr;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, ctx->stream);
    return e;
  }
  void end(hipEvent_t e0, int kind, double points) {
    if (!e0) return;
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return;
    }
    (void)hipEventRecord(e1, ctx->stream);
    recs.push_back({e0, e1, kind, points});
  }
  void harvest() {
    for (auto &r : recs) {
      float ms = 0;
      if (hipEventSynchronize(r.e1) == hipSuccess && hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
        if (r.kind == 0) {
          ctx->fill_prof[0] += ms;
          ctx->fill_prof[1] += 1;
          ctx->fill_prof[2] += r.points;
        } else if (r.kind == 1) {
          ctx->fill_prof[3] += ms;
          ctx->fill_prof[4] += 1;
        } else {
          ctx->fill_prof[5] += ms;
        }
      }
      (void)hipEventDestroy(r.e0);
      (void)hipEventDestroy(r.e1);
    }
    recs.clear();
  }
  ~FillTimer() { harvest(); }
};

// test hook: the lossless variant of the pair kernel (mom_test_fill_lossless)
static int g_fill_lossless = 1;

int fill_domain(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                const FillTarget &t, const Image *img) {
  int rc = compute_quad_points(ctx, m, rule);
  if (rc) return rc;
  const cd I(0.0, 1.0);
  const cd k = ri * omega / kC0;
  // sysmat.f90:301 and 388: c1 = -i omega mu0, c2 = -1/(i omega eps0 ri^2)
  const cd c1 = -(I * omega) * mu0();
  const cd c2 = -1.0 / ((I * omega) * kEps0 * (ri * ri));
  FarArgs a;
  a.F = (int)m->F;
  a.Q = rule.Q;
  for (int i = 0; i < kMaxQ; i++) a.w[i] = i < rule.Q ? rule.w[i] : 0.0;
  a.qp = ctx->qp.as<double>();
  a.cp = m->fcp.as<double>();
  a.ccls = m->cls();
  a.area = m->farea.as<double>();
  const bool glob = t.mode == 0 && !t.local;  // global unknown numbering
  a.side = sides(m, glob);
  a.kr = k.real();
  a.ki = k.imag();
  // point distances: |r - r'| <= extent; with an image |J r - r'| <= 2 rmax
  const double reach = img ? std::max(m->extent, 2.0 * m->rmax) : m->extent;
  if (!(std::fabs(a.kr) * reach < kCisMaxArg))
    return fail(MOM_ERR_ARG, "Re(k) times the largest point distance is beyond the fill's phase range (2^19 rad)");
  const double thr = m->avelen * 3;
  a.thr2 = thr * thr;
  a.c1 = make_double2(c1.real(), c1.imag());
  a.c2 = make_double2(c2.real(), c2.imag());
  a.image = img ? 1 : 0;
  a.gen = (img && !img->diag) ? 1 : 0;
  for (int c = 0; c < 9; c++) {
    a.M[c] = img ? img->M[c] : (c % 4 == 0 ? 1.0 : 0.0);
    a.J[c] = img ? img->M[3 * (c % 3) + c / 3] : (c % 4 == 0 ? 1.0 : 0.0);
  }
  for (int c = 0; c < 3; c++) a.js[c] = a.M[4 * c];
  // the transposed blocks of an image belong to its inverse element: an
  // image that is not its own inverse (a rotation by other than pi)
  // integrates every ordered pair and has no transposed assembly
  const bool all_pairs = img && !img->invol;
  a.out = t;
  // far/smooth pairs, slab by slab of source faces, through the face-side
  // planes (no atomics on the matrix)
  const int S = (int)std::min<int64_t>(m->F, kSlab);
  const size_t bbytes = sizeof(double2) * 18 * (size_t)S * m->F;
  if (ctx->scratch.bytes < bbytes && (rc = ctx->scratch.alloc(bbytes))) return rc;
  // output row -> local edge
  const int64_t nrows = glob ? t.Noff : m->E;
  std::vector<int> rl(nrows, -1);
  std::vector<int> colidx(m->E);
  for (int64_t e = 0; e < m->E; e++) {
    const int g = glob ? m->h_eparent[e] : (int)e;
    rl[g] = (int)e;
    colidx[e] = g;
  }
  // local edges by their highest face, descending: the source edges with a
  // side beyond slab k form a prefix
  std::vector<int> hside(2 * m->E);
  MOM_TRY(copy_d2h(ctx, hside.data(), m->eside.ptr, sizeof(int) * 2 * m->E));
  std::vector<int> maxface(m->E), bydesc(m->E);
  for (int64_t e = 0; e < m->E; e++) {
    maxface[e] = std::max(hside[2 * e] < 0 ? -1 : hside[2 * e] / 3, hside[2 * e + 1] < 0 ? -1 : hside[2 * e + 1] / 3);
    bydesc[e] = (int)e;
  }
  std::stable_sort(bydesc.begin(), bydesc.end(), [&](int x, int y) { return maxface[x] > maxface[y]; });
  DevBuf drl, dcol, ddesc;
  if ((rc = drl.alloc(sizeof(int) * nrows)) || (rc = dcol.alloc(sizeof(int) * m->E)) ||
      (rc = ddesc.alloc(sizeof(int) * m->E)))
    return rc;
  MOM_TRY(copy_h2d(ctx, drl.ptr, rl.data(), sizeof(int) * nrows));
  MOM_TRY(copy_h2d(ctx, dcol.ptr, colidx.data(), sizeof(int) * m->E));
  MOM_TRY(copy_h2d(ctx, ddesc.ptr, bydesc.data(), sizeof(int) * m->E));
  PlaneArgs pa;
  pa.f = a;
  pa.B = ctx->scratch.as<double2>();
  pa.S = S;
  pa.mode = t.mode;
  AssembleArgs ga;
  ga.B = pa.B;
  ga.F = (int)m->F;
  ga.S = S;
  ga.row_local = drl.as<int>();
  ga.nrows = nrows;
  ga.eside = m->eside.as<int>();
  ga.col_index = dcol.as<int>();
  ga.ktr = img ? img->detj : 1.0;
  ga.out = t;
  const int nslab = (int)((m->F + S - 1) / S);
  FillTimer timer(ctx);
  for (int k = 0; k < nslab; k++) {
    pa.n0 = k * S;
    // pairs with both faces below n0 were integrated by earlier slabs
    pa.mlo = all_pairs ? 0 : pa.n0;
    const int scur = (int)std::min<int64_t>(S, m->F - pa.n0);
    auto grid_of = [&](int spw) {
      return dim3((unsigned)((m->F - pa.mlo + 63) / 64), (unsigned)((scur + kWaves * spw - 1) / (kWaves * spw)));
    };
    // a lossless medium (the vacuum outside) skips the exponential of e^{ikR}
    const bool ll = a.ki == 0.0 && g_fill_lossless;
    auto launch13 = [&](int spw) {
      if (spw == 2) {
        if (ll)
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 2, true>), grid_of(2), dim3(256), 0, ctx->stream, pa);
        else
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 2>), grid_of(2), dim3(256), 0, ctx->stream, pa);
      } else if (spw == 8) {
        if (ll)
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 8, true>), grid_of(8), dim3(256), 0, ctx->stream, pa);
        else
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 8>), grid_of(8), dim3(256), 0, ctx->stream, pa);
      } else {
        if (ll)
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 4, true>), grid_of(4), dim3(256), 0, ctx->stream, pa);
        else
          hipLaunchKernelGGL((pair_plane_kernel<13, false, 4>), grid_of(4), dim3(256), 0, ctx->stream, pa);
      }
    };
    if (t.dk) {
      hipLaunchKernelGGL((pair_plane_kernel<0, true>), grid_of(4), dim3(256), 0, ctx->stream, pa);
    } else if (rule.Q == 13) {
      if (ctx->fill_spw == 0 && k == 0) {
        // autotune on this slab: each candidate rewrites the same planes
        hipEvent_t e0, e1;
        MOM_HIP(hipEventCreate(&e0));
        MOM_HIP(hipEventCreate(&e1));
        const int cand[3] = {2, 4, 8};
        float best = 1e30f;
        for (int c = 0; c < 3; c++) {
          MOM_HIP(hipEventRecord(e0, ctx->stream));
          launch13(cand[c]);
          MOM_HIP(hipGetLastError());
          MOM_HIP(hipEventRecord(e1, ctx->stream));
          MOM_HIP(hipEventSynchronize(e1));
          float ms = 0;
          MOM_HIP(hipEventElapsedTime(&ms, e0, e1));
          ctx->fill_spw_ms[c] = ms;
          if (ms < best) {
            best = ms;
            ctx->fill_spw = cand[c];
          }
        }
        MOM_HIP(hipEventDestroy(e0));
        MOM_HIP(hipEventDestroy(e1));
      }
      hipEvent_t ev = timer.begin();
      launch13(ctx->fill_spw ? ctx->fill_spw : 4);
      timer.end(ev, 0, (double)(m->F - pa.mlo) * scur * rule.Q * rule.Q);
    } else {
      hipLaunchKernelGGL(pair_plane_kernel<0>, grid_of(4), dim3(256), 0, ctx->stream, pa);
    }
    MOM_HIP(hipGetLastError());
    const int ne = m->h_slab_ptr[k + 1] - m->h_slab_ptr[k];
    if (ne > 0) {
      ga.n0 = pa.n0;
      ga.mlo = pa.mlo;
      ga.slab_edges = m->slab_edges.as<int>() + m->h_slab_ptr[k];
      hipEvent_t ev = timer.begin();
      hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)((nrows + 255) / 256), (unsigned)ne), dim3(256), 0,
                         ctx->stream, ga);
      MOM_HIP(hipGetLastError());
      // transposed blocks: test sides in the slab, source sides beyond it
      const int beyond = pa.n0 + scur;
      int nb = 0;
      while (!all_pairs && nb < (int)m->E && maxface[bydesc[nb]] >= beyond) nb++;
      if (nb > 0) {
        hipLaunchKernelGGL(assemble_transposed_kernel, dim3((unsigned)((ne + kTR - 1) / kTR), (unsigned)((nb + kTC - 1) / kTC)),
                           dim3(256), 0, ctx->stream, ga, ddesc.as<int>(), ne, nb);
        MOM_HIP(hipGetLastError());
      }
      timer.end(ev, 1, 0.0);
    }
  }
  // the row/column maps must outlive the queued kernels
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  timer.harvest();

  // near pairs of the image: the reflected test faces against the sources
  DevBuf inear, ioff;
  int64_t nnear = m->nnear;
  const int *pairs = m->near.as<int>();
  const int64_t *off = m->near_off.as<int64_t>();
  if (img) {
    std::vector<int32_t> np;
    near_pairs(m->h_cls(), m->F, m->avelen, np, a.J);
    nnear = (int64_t)np.size() / 2;
    if (nnear > 0) {
      std::vector<int64_t> ho;
      near_offsets(np, m->F, ho);
      if ((rc = inear.alloc(sizeof(int32_t) * np.size())) || (rc = ioff.alloc(sizeof(int64_t) * ho.size())))
        return rc;
      MOM_TRY(copy_h2d(ctx, inear.ptr, np.data(), sizeof(int32_t) * np.size()));
      MOM_TRY(copy_h2d(ctx, ioff.ptr, ho.data(), sizeof(int64_t) * ho.size()));
    }
    pairs = inear.as<int>();
    off = ioff.as<int64_t>();
  }
  if (nnear > 0) {
    // per pair blocks into scratch (the slab planes are assembled by now),
    // then one fixed-order sum per matrix entry
    const size_t nbytes = sizeof(double2) * 18 * (size_t)nnear;
    if (ctx->scratch.bytes < nbytes && (rc = ctx->scratch.alloc(nbytes))) return rc;
    NearArgs b;
    b.Q = rule.Q;
    for (int i = 0; i < kMaxQ; i++) b.w[i] = a.w[i];
    b.pairs = pairs;
    b.npairs = nnear;
    b.qp = a.qp;
    b.p = m->p.as<double>();
    b.fn = m->fn.as<double>();
    b.fm = m->fm.as<double>();
    b.fs = m->fs.as<double>();
    b.area = a.area;
    b.fnodes = m->fnodes.as<int>();
    b.side = a.side;
    b.mode = t.mode;
    // analytic part [1/R - k^2 R/2]/4pi; with dk its k-derivative -k R/4pi
    b.kk4 = make_double2(t.dk ? 0.0 : 1.0 / (4.0 * MOM_PI), 0.0);
    const cd kk8 = t.dk ? k / (4.0 * MOM_PI) : (k * k) / (8.0 * MOM_PI);
    b.kk8 = make_double2(kk8.real(), kk8.imag());
    b.c1 = a.c1;
    b.c2 = a.c2;
    for (int c = 0; c < 3; c++) b.js[c] = a.js[c];
    b.image = a.image;
    b.gen = a.gen;
    for (int c = 0; c < 9; c++) b.M[c] = a.M[c];
    b.nb = ctx->scratch.as<double2>();
    hipEvent_t ev = timer.begin();
    hipLaunchKernelGGL(near_fill_kernel, dim3((unsigned)((nnear + 127) / 128)), dim3(128), 0,
                       ctx->stream, b);
    MOM_HIP(hipGetLastError());
    NearEmitArgs e;
    e.pairs = pairs;
    e.off = off;
    e.npairs = nnear;
    e.eside = m->eside.as<int>();
    e.s_l = m->s_l.as<int>();
    e.side = a.side;
    e.nb = b.nb;
    e.mode = t.mode;
    e.out = t;
    hipLaunchKernelGGL(near_emit_kernel, dim3((unsigned)((9 * nnear + 255) / 256)), dim3(256), 0, ctx->stream, e);
    MOM_HIP(hipGetLastError());
    timer.end(ev, 2, 0.0);
    // the image's pair list must outlive the kernels
    if (img) MOM_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (!timer.recs.empty()) timer.harvest();
  return 0;
}

// ---------------------------------------------------------------------------
// Plane-wave right-hand sides (source.f90:136-247 with pw_fields 84-101):
// one thread per (face, source, representation); each face adds
// <f_p, E_inc> and <f_p, H_inc> of its three sides.  For a symmetry group
// the incident fields are projected on the representation first
// (src_field_frags, source.f90:249-315): E_r(r) = (1/n) sum_g ef[g][r]
// J_g E(J_g r) and H_r likewise with ef det J_g.
struct PwSrc {
  double dir[3], pol[3];
};

// the plane wave's E and H = k/(omega mu0) dir x E at r (pw_fields)
__device__ __forceinline__ void pw_eh(const PwSrc &S, const double *r, double kr, double ki, double2 hfac,
                                      double E[3][2], double H[3][2]) {
  double kd = S.dir[0] * r[0] + S.dir[1] * r[1] + S.dir[2] * r[2];
  // exp(i k d.r)
  double ex = exp(-ki * kd), sn, cs;
  sincos(kr * kd, &sn, &cs);
  double er = ex * cs, ei = ex * sn;
  for (int c = 0; c < 3; c++) {
    E[c][0] = S.pol[c] * er;
    E[c][1] = S.pol[c] * ei;
  }
  for (int c = 0; c < 3; c++) {
    int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    double xr = S.dir[c1] * E[c2][0] - S.dir[c2] * E[c1][0];
    double xi = S.dir[c1] * E[c2][1] - S.dir[c2] * E[c1][1];
    H[c][0] = hfac.x * xr - hfac.y * xi;
    H[c][1] = hfac.x * xi + hfac.y * xr;
  }
}

// y = J x and y = J^T x for a row-major 3 x 3 J (MATMUL order)
__device__ __forceinline__ void jac_mul(const double *J, const double *x, double *y) {
  for (int i = 0; i < 3; i++) y[i] = J[3 * i] * x[0] + J[3 * i + 1] * x[1] + J[3 * i + 2] * x[2];
}
__device__ __forceinline__ void jac_tmul(const double *J, const double *x, double *y) {
  for (int i = 0; i < 3; i++) y[i] = J[i] * x[0] + J[3 + i] * x[1] + J[6 + i] * x[2];
}

// w[(col 3F + 3 f + p) 2 + {0, 1}]: <f_p, E_r>, <f_p, H_r> of face side p of
// face f for column col = source * G.n + representation.  With a group
// (src_field_frags, source.f90:250-315): E_r(r) = (1/n) sum_g ef[g][r] J_g
// E(J_g^T r), H_r the same with det J_g.
__global__ void srcvec_kernel(int F, int Q, QuadRule rule, const double *__restrict__ qp,
                              const double *__restrict__ area, FaceSide side, int nsrc,
                              const PwSrc *__restrict__ src, double kr, double ki, double2 hfac,
                              double2 *__restrict__ w, Group G) {
  int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)F * nsrc * G.n) return;
  const int f = (int)(id % F);
  const int64_t sr = id / F;
  const int rep = (int)(sr % G.n), s = (int)(sr / G.n);
  const PwSrc S = src[s];
  double Ie[3][2] = {{0, 0}, {0, 0}, {0, 0}}, Ih[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int t = 0; t < Q; t++) {
    const double *r = qp + ((int64_t)f * Q + t) * 3;
    double E[3][2], H[3][2];
    if (G.n == 1) {
      pw_eh(S, r, kr, ki, hfac, E, H);
    } else {
      for (int c = 0; c < 3; c++) E[c][0] = E[c][1] = H[c][0] = H[c][1] = 0.0;
      for (int g = 0; g < G.n; g++) {
        double pt[3];
        jac_tmul(G.J[g], r, pt);
        double Eg[3][2], Hg[3][2];
        pw_eh(S, pt, kr, ki, hfac, Eg, Hg);
        const double2 e = G.ef[g][rep], eh = make_double2(e.x * G.detj[g], e.y * G.detj[g]);
        double Er[3], Ei[3], Hr[3], Hi[3];
        const double er[3] = {Eg[0][0], Eg[1][0], Eg[2][0]}, ei[3] = {Eg[0][1], Eg[1][1], Eg[2][1]};
        const double hr[3] = {Hg[0][0], Hg[1][0], Hg[2][0]}, hi[3] = {Hg[0][1], Hg[1][1], Hg[2][1]};
        jac_mul(G.J[g], er, Er);
        jac_mul(G.J[g], ei, Ei);
        jac_mul(G.J[g], hr, Hr);
        jac_mul(G.J[g], hi, Hi);
        for (int c = 0; c < 3; c++) {
          E[c][0] += Er[c] * e.x - Ei[c] * e.y;
          E[c][1] += Ei[c] * e.x + Er[c] * e.y;
          H[c][0] += Hr[c] * eh.x - Hi[c] * eh.y;
          H[c][1] += Hi[c] * eh.x + Hr[c] * eh.y;
        }
      }
      for (int c = 0; c < 3; c++) {
        E[c][0] /= G.n;
        E[c][1] /= G.n;
        H[c][0] /= G.n;
        H[c][1] /= G.n;
      }
    }
    for (int p = 0; p < 3; p++) {
      const double cf = side.coef[3 * f + p];
      const double *v = side.v + 3 * (3 * f + p);
      double fv[3] = {cf * (r[0] - v[0]), cf * (r[1] - v[1]), cf * (r[2] - v[2])};
      double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
      for (int c = 0; c < 3; c++) {
        a0 += fv[c] * E[c][0];
        a1 += fv[c] * E[c][1];
        b0 += fv[c] * H[c][0];
        b1 += fv[c] * H[c][1];
      }
      Ie[p][0] += rule.w[t] * a0;
      Ie[p][1] += rule.w[t] * a1;
      Ih[p][0] += rule.w[t] * b0;
      Ih[p][1] += rule.w[t] * b1;
    }
  }
  const double Am = area[f];
  double2 *ws = w + 2 * (sr * 3 * (int64_t)F + 3 * f);
  for (int p = 0; p < 3; p++) {
    ws[2 * p] = make_double2(Ie[p][0] * Am, Ie[p][1] * Am);
    ws[2 * p + 1] = make_double2(Ih[p][0] * Am, Ih[p][1] * Am);
  }
}

// q[g, col] += the edge's face sides in face order (srcvec's loop over the
// faces, source.f90:198-244), one thread per (local edge, column): no
// atomics, so the sums are the same on every run.
__global__ void srcvec_gather_kernel(int64_t E, int F, int64_t ncol, const int *__restrict__ eside,
                                     const int *__restrict__ s_g, const double2 *__restrict__ w, int64_t Nt,
                                     double2 *q, int64_t ldq) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= E * ncol) return;
  const int64_t e = id % E, col = id / E;
  int c0 = eside[2 * e], c1 = eside[2 * e + 1];
  if (c0 < 0 || (c1 >= 0 && c1 / 3 < c0 / 3)) {
    const int t = c0;
    c0 = c1;
    c1 = t;
  }
  if (c0 < 0) return;
  const double2 *wc = w + 2 * col * 3 * (int64_t)F;
  double2 se = wc[2 * c0], sh = wc[2 * c0 + 1];
  if (c1 >= 0) {
    se.x += wc[2 * c1].x;
    se.y += wc[2 * c1].y;
    sh.x += wc[2 * c1 + 1].x;
    sh.y += wc[2 * c1 + 1].y;
  }
  const int64_t g = s_g[c0];
  double2 *qs = q + col * ldq;
  qs[g] = make_double2(qs[g].x + se.x, qs[g].y + se.y);
  qs[g + Nt] = make_double2(qs[g + Nt].x + sh.x, qs[g + Nt].y + sh.y);
}

int srcvec_pw(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
              int64_t nsrc, const double *angles, int64_t Nt, double2 *q, int64_t ldq, const Group &G) {
  int rc = compute_quad_points(ctx, m, rule);
  if (rc) return rc;
  std::vector<PwSrc> hs(nsrc);
  for (int64_t s = 0; s < nsrc; s++) {
    double th = angles[3 * s], ph = angles[3 * s + 1], ps = angles[3 * s + 2];
    // aux.f90:245-261 get_dir / get_pol
    hs[s].dir[0] = sin(th) * cos(ph);
    hs[s].dir[1] = sin(th) * sin(ph);
    hs[s].dir[2] = cos(th);
    hs[s].pol[0] = cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph);
    hs[s].pol[1] = -cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph);
    hs[s].pol[2] = sin(ps) * sin(th);
  }
  const int64_t ncol = nsrc * G.n;
  DevBuf dsrc, dw;
  if ((rc = dsrc.alloc(sizeof(PwSrc) * nsrc)) || (rc = dw.alloc(sizeof(double2) * 6 * m->F * ncol))) return rc;
  MOM_TRY(copy_h2d(ctx, dsrc.ptr, hs.data(), sizeof(PwSrc) * nsrc));
  const cd k = ri * omega / kC0;
  const cd hfac = k / (omega * mu0());
  int64_t total = m->F * ncol;
  hipLaunchKernelGGL(srcvec_kernel, dim3((unsigned)((total + 127) / 128)), dim3(128), 0, ctx->stream,
                     (int)m->F, rule.Q, rule, ctx->qp.as<double>(), m->farea.as<double>(),
                     sides(m, true), (int)nsrc, dsrc.as<PwSrc>(), k.real(), k.imag(),
                     make_double2(hfac.real(), hfac.imag()), dw.as<double2>(), G);
  hipLaunchKernelGGL(srcvec_gather_kernel, dim3((unsigned)((m->E * ncol + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->E, (int)m->F, ncol, m->eside.as<int>(), m->s_g.as<int>(), dw.as<double2>(), Nt, q, ldq);
  MOM_HIP(hipGetLastError());
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// cs_prtsrf (cs.f90:14-104): per face, the quadrature sums of the total
// (absorption) and scattered Poynting flux; the host adds the per-face
// values in face order as the reference does.  With a symmetry group every
// image g of the face contributes: normal J n, tangential fields
// sum_r ef[g][r] J (n x f) x_r (with det J for H) from the representations'
// solutions, incident field at J r.
__global__ void cs_kernel(int F, int Q, QuadRule rule, const double *__restrict__ qp,
                          const double *__restrict__ area, const double *__restrict__ fnrm,
                          FaceSide side, const double2 *__restrict__ x, int64_t Nt, PwSrc S,
                          double kr, double ki, double2 hfac, double *out, Group G, int64_t ldx) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const double *nor0 = fnrm + 3 * f;
  double int1 = 0, int2 = 0;
  for (int t = 0; t < Q; t++) {
    const double *r0 = qp + ((int64_t)f * Q + t) * 3;
   for (int gi = 0; gi < G.n; gi++) {
    const double *Jg = G.J[gi];
    const bool dg = Jg[1] == 0.0 && Jg[2] == 0.0 && Jg[3] == 0.0 && Jg[5] == 0.0 && Jg[6] == 0.0 && Jg[7] == 0.0;
    double nor[3], r[3];
    jac_mul(Jg, nor0, nor);
    jac_mul(Jg, r0, r);
    double Et[3][2] = {{0, 0}, {0, 0}, {0, 0}}, Ht[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int rep = 0; rep < G.n; rep++) {
      const double2 e = G.ef[gi][rep], eh = make_double2(e.x * G.detj[gi], e.y * G.detj[gi]);
      const double2 *xr = x + rep * ldx;
      for (int p = 0; p < 3; p++) {
        const double cf = side.coef[3 * f + p];
        const double *v = side.v + 3 * (3 * f + p);
        double fv[3] = {cf * (r0[0] - v[0]), cf * (r0[1] - v[1]), cf * (r0[2] - v[2])};
        double fn0[3] = {nor0[1] * fv[2] - nor0[2] * fv[1], nor0[2] * fv[0] - nor0[0] * fv[2],
                         nor0[0] * fv[1] - nor0[1] * fv[0]};
        double fn[3];
        if (dg)
          for (int c = 0; c < 3; c++) fn[c] = fn0[c] * Jg[4 * c];
        else
          jac_mul(Jg, fn0, fn);
        const int g = side.g[3 * f + p];
        const double2 xj = xr[g], xm = xr[g + Nt];
        if (e.y == 0.0) {
          for (int c = 0; c < 3; c++) {
            Ht[c][0] -= xj.x * fn[c] * eh.x;
            Ht[c][1] -= xj.y * fn[c] * eh.x;
            Et[c][0] += xm.x * fn[c] * e.x;
            Et[c][1] += xm.y * fn[c] * e.x;
          }
        } else {
          const double2 je = make_double2(xj.x * eh.x - xj.y * eh.y, xj.x * eh.y + xj.y * eh.x);
          const double2 me = make_double2(xm.x * e.x - xm.y * e.y, xm.x * e.y + xm.y * e.x);
          for (int c = 0; c < 3; c++) {
            Ht[c][0] -= je.x * fn[c];
            Ht[c][1] -= je.y * fn[c];
            Et[c][0] += me.x * fn[c];
            Et[c][1] += me.y * fn[c];
          }
        }
      }
    }
    double Ei[3][2], Hi[3][2];
    pw_eh(S, r, kr, ki, hfac, Ei, Hi);
    // tangential incident fields (n x F) x n = F - (n.F) n
    double nE[2] = {0, 0}, nH[2] = {0, 0};
    for (int c = 0; c < 3; c++) {
      nE[0] += nor[c] * Ei[c][0];
      nE[1] += nor[c] * Ei[c][1];
      nH[0] += nor[c] * Hi[c][0];
      nH[1] += nor[c] * Hi[c][1];
    }
    double Es[3][2], Hs[3][2];
    for (int c = 0; c < 3; c++) {
      Es[c][0] = Et[c][0] - (Ei[c][0] - nE[0] * nor[c]);
      Es[c][1] = Et[c][1] - (Ei[c][1] - nE[1] * nor[c]);
      Hs[c][0] = Ht[c][0] - (Hi[c][0] - nH[0] * nor[c]);
      Hs[c][1] = Ht[c][1] - (Hi[c][1] - nH[1] * nor[c]);
    }
    // 0.5 Re[(E x conj H) . n]
    double s1 = 0, s2 = 0;
    for (int c = 0; c < 3; c++) {
      int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      double re1 = (Et[c1][0] * Ht[c2][0] + Et[c1][1] * Ht[c2][1]) -
                   (Et[c2][0] * Ht[c1][0] + Et[c2][1] * Ht[c1][1]);
      double re2 = (Es[c1][0] * Hs[c2][0] + Es[c1][1] * Hs[c2][1]) -
                   (Es[c2][0] * Hs[c1][0] + Es[c2][1] * Hs[c1][1]);
      s1 += re1 * nor[c];
      s2 += re2 * nor[c];
    }
    int1 += rule.w[t] * 0.5 * s1;
    int2 += rule.w[t] * 0.5 * s2;
   }
  }
  out[2 * f] = int1 * area[f];
  out[2 * f + 1] = int2 * area[f];
}

int cs_partials(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                const double2 *x, int64_t Nt, double th, double ph, double ps, double *partial, const Group &G,
                int64_t ldx) {
  int rc = compute_quad_points(ctx, m, rule);
  if (rc) return rc;
  PwSrc S;
  S.dir[0] = sin(th) * cos(ph);
  S.dir[1] = sin(th) * sin(ph);
  S.dir[2] = cos(th);
  S.pol[0] = cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph);
  S.pol[1] = -cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph);
  S.pol[2] = sin(ps) * sin(th);
  const cd k = ri * omega / kC0;
  const cd hfac = k / (omega * mu0());
  hipLaunchKernelGGL(cs_kernel, dim3((unsigned)((m->F + 127) / 128)), dim3(128), 0, ctx->stream,
                     (int)m->F, rule.Q, rule, ctx->qp.as<double>(), m->farea.as<double>(),
                     m->fn.as<double>(), sides(m, true), x, Nt, S, k.real(), k.imag(),
                     make_double2(hfac.real(), hfac.imag()), partial, G, ldx > 0 ? ldx : 2 * Nt);
  MOM_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Far fields (ffields.f90:16-122, identity group) for many directions at
// once.  The direction-independent part of far_fields, the surface currents
// at the quadrature points weighted by w_t A_n, is formed once per point:
// J(s) = w_t A_n sum_q x[g_q] f_q(r_t) and M(s) likewise with x[g_q + N].
// Per direction the four radiation integrals are then one pass over the
// points: Nt, Np = sum_s phasor(s) (u_theta, u_phi) . J(s), Lt, Lp from M,
// phasor = exp(-i k dir . r_s).  The reference sums edge by edge within a
// face; here the three edges of a face are summed at each point first, a
// difference at the level of rounding.
// With a symmetry group the points are those of every image g of the
// surface (J r) with currents J f sum_r ef[g][r] x_r (det J for M), as
// ffields.f90:67-94 sums them.
__global__ void farfield_sources_kernel(int F, int Q, QuadRule rule, const double *__restrict__ qp,
                                        const double *__restrict__ area, FaceSide side,
                                        const double2 *__restrict__ x, int64_t Nt, double *__restrict__ pts,
                                        double2 *__restrict__ cur, int64_t S, Group G, int64_t ldx) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int64_t FQ = (int64_t)F * Q;
  const int gi = (int)(s / FQ);
  const int64_t s0 = s - gi * FQ;
  const int f = (int)(s0 / Q), t = (int)(s0 - (int64_t)f * Q);
  const double *Jg = G.J[gi];
  const bool dg = Jg[1] == 0.0 && Jg[2] == 0.0 && Jg[3] == 0.0 && Jg[5] == 0.0 && Jg[6] == 0.0 && Jg[7] == 0.0;
  const double *r = qp + s0 * 3;
  double J[3][2] = {{0, 0}, {0, 0}, {0, 0}}, M[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int rep = 0; rep < G.n; rep++) {
    const double2 e = G.ef[gi][rep], em = make_double2(e.x * G.detj[gi], e.y * G.detj[gi]);
    const double2 *xr = x + rep * ldx;
    for (int p = 0; p < 3; p++) {
      const int g = side.g[3 * f + p];
      if (g < 0) continue;
      const double cf = side.coef[3 * f + p];
      const double *v = side.v + 3 * (3 * f + p);
      const double f0[3] = {cf * (r[0] - v[0]), cf * (r[1] - v[1]), cf * (r[2] - v[2])};
      double fv[3];
      if (dg)
        for (int c = 0; c < 3; c++) fv[c] = Jg[4 * c] * f0[c];
      else
        jac_mul(Jg, f0, fv);
      const double2 xj = xr[g], xm = xr[g + Nt];
      if (e.y == 0.0) {
        for (int c = 0; c < 3; c++) {
          J[c][0] += xj.x * e.x * fv[c];
          J[c][1] += xj.y * e.x * fv[c];
          M[c][0] += xm.x * em.x * fv[c];
          M[c][1] += xm.y * em.x * fv[c];
        }
      } else {
        const double2 je = make_double2(xj.x * e.x - xj.y * e.y, xj.x * e.y + xj.y * e.x);
        const double2 me = make_double2(xm.x * em.x - xm.y * em.y, xm.x * em.y + xm.y * em.x);
        for (int c = 0; c < 3; c++) {
          J[c][0] += je.x * fv[c];
          J[c][1] += je.y * fv[c];
          M[c][0] += me.x * fv[c];
          M[c][1] += me.y * fv[c];
        }
      }
    }
  }
  const double wa = rule.w[t] * area[f];
  double pr[3];
  if (dg)
    for (int c = 0; c < 3; c++) pr[c] = Jg[4 * c] * r[c];
  else
    jac_mul(Jg, r, pr);
  for (int c = 0; c < 3; c++) {
    pts[3 * s + c] = pr[c];
    cur[c * S + s] = make_double2(wa * J[c][0], wa * J[c][1]);
    cur[(3 + c) * S + s] = make_double2(wa * M[c][0], wa * M[c][1]);
  }
}

constexpr int kFfTile = 256, kFfChunk = 16 * kFfTile;  // points per LDS tile and per partial sum

// One thread per direction; blockIdx.y takes kFfChunk points, staged through
// LDS a tile at a time (every thread reads the same point: broadcast).
// dirs[8 d ..]: dir (st cp, st sp, ct), u_theta (ct cp, ct sp, -st), u_phi
// (-sp, cp).  part[chunk][d][4] = (Nt, Np, Lt, Lp) over the chunk.
__global__ __launch_bounds__(256) void farfield_dir_kernel(int64_t S, const double *__restrict__ pts,
                                                            const double2 *__restrict__ cur, int ndir,
                                                            const double *__restrict__ dirs, double kr,
                                                            double ki, double2 *__restrict__ part) {
  __shared__ double sp[3][kFfTile];
  __shared__ double2 sc[6][kFfTile];
  const int tid = threadIdx.x, d = blockIdx.x * 256 + tid;
  const int64_t s0 = (int64_t)blockIdx.y * kFfChunk, s1 = std::min<int64_t>(S, s0 + kFfChunk);
  double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (d < ndir)
    for (int c = 0; c < 8; c++) u[c] = dirs[8 * d + c];
  double a[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int64_t tile = s0; tile < s1; tile += kFfTile) {
    const int64_t i = tile + tid;
    if (i < s1) {
      for (int c = 0; c < 3; c++) sp[c][tid] = pts[3 * i + c];
      for (int c = 0; c < 6; c++) sc[c][tid] = cur[c * S + i];
    }
    __syncthreads();
    const int nt = (int)std::min<int64_t>(kFfTile, s1 - tile);
    for (int q = 0; q < nt; q++) {
      const double rd = u[0] * sp[0][q] + u[1] * sp[1][q] + u[2] * sp[2][q];
      double er, ei;
      cis_exp(ki * rd, -kr * rd, er, ei);  // exp(-i k rd)
      double2 pr[4];
      pr[0] = make_double2(u[3] * sc[0][q].x + u[4] * sc[1][q].x + u[5] * sc[2][q].x,
                           u[3] * sc[0][q].y + u[4] * sc[1][q].y + u[5] * sc[2][q].y);
      pr[1] = make_double2(u[6] * sc[0][q].x + u[7] * sc[1][q].x, u[6] * sc[0][q].y + u[7] * sc[1][q].y);
      pr[2] = make_double2(u[3] * sc[3][q].x + u[4] * sc[4][q].x + u[5] * sc[5][q].x,
                           u[3] * sc[3][q].y + u[4] * sc[4][q].y + u[5] * sc[5][q].y);
      pr[3] = make_double2(u[6] * sc[3][q].x + u[7] * sc[4][q].x, u[6] * sc[3][q].y + u[7] * sc[4][q].y);
      for (int k = 0; k < 4; k++) {
        a[k][0] += pr[k].x * er - pr[k].y * ei;
        a[k][1] += pr[k].x * ei + pr[k].y * er;
      }
    }
    __syncthreads();
  }
  if (d < ndir)
    for (int k = 0; k < 4; k++) part[((int64_t)blockIdx.y * ndir + d) * 4 + k] = make_double2(a[k][0], a[k][1]);
}

// out[d][4] = sum over the chunks, in chunk order
__global__ void farfield_reduce_kernel(int ndir, int nchunk, const double2 *__restrict__ part, double2 *out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndir) return;
  for (int k = 0; k < 4; k++) {
    double re = 0, im = 0;
    for (int c = 0; c < nchunk; c++) {
      const double2 v = part[((int64_t)c * ndir + d) * 4 + k];
      re += v.x;
      im += v.y;
    }
    out[4 * d + k] = make_double2(re, im);
  }
}

// Direction frames of far_fields (ffields.f90:30-45): dirs[8 d ..] = dir
// (st cp, st sp, ct), u_theta (ct cp, ct sp, -st), u_phi (-sp, cp).
__global__ void farfield_dirs_kernel(int64_t ndir, const double *__restrict__ theta, const double *__restrict__ phi,
                                     double *__restrict__ dirs) {
  const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (d >= ndir) return;
  double st, ct, sp, cp;
  sincos(theta[d], &st, &ct);
  sincos(phi[d], &sp, &cp);
  double *u = dirs + 8 * d;
  u[0] = st * cp;
  u[1] = st * sp;
  u[2] = ct;
  u[3] = ct * cp;
  u[4] = ct * sp;
  u[5] = -st;
  u[6] = -sp;
  u[7] = cp;
}

// ffields.f90:107-113: with v = (Nt, Np, Lt, Lp) of direction d,
// et = P (Lp + a Nt), ep = -P (Lt - a Np), ht = -P (Np - Lt b),
// hp = P (Nt + Lp b), P = i k exp(ikr) / (4 pi r), a = eta0 / ri,
// b = ri / eta0; out[8 d ..] = et, ep, ht, hp (interleaved complex).
struct FarCoef {
  double2 P, a, b;
};

#pragma clang fp contract(off)
__device__ __forceinline__ double2 zmul(double2 u, double2 v) {
  return make_double2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
}
__device__ __forceinline__ double2 zadd(double2 u, double2 v) { return make_double2(u.x + v.x, u.y + v.y); }
__device__ __forceinline__ double2 zsub(double2 u, double2 v) { return make_double2(u.x - v.x, u.y - v.y); }
__device__ __forceinline__ double2 zneg(double2 u) { return make_double2(-u.x, -u.y); }

__global__ void farfield_combine_kernel(int64_t ndir, const double2 *__restrict__ v, FarCoef c,
                                        double *__restrict__ out) {
  const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (d >= ndir) return;
  const double2 Nt = v[4 * d], Np = v[4 * d + 1], Lt = v[4 * d + 2], Lp = v[4 * d + 3];
  const double2 f[4] = {zmul(c.P, zadd(Lp, zmul(c.a, Nt))), zneg(zmul(c.P, zsub(Lt, zmul(c.a, Np)))),
                        zneg(zmul(c.P, zsub(Np, zmul(Lt, c.b)))), zmul(c.P, zadd(Nt, zmul(Lp, c.b)))};
  for (int k = 0; k < 4; k++) {
    out[8 * d + 2 * k] = f[k].x;
    out[8 * d + 2 * k + 1] = f[k].y;
  }
}
#pragma clang fp contract(on)

// Device work areas of one far-field evaluation, laid out in one block of
// `far_work_bytes` bytes (each part 256-byte aligned).
struct FarWork {
  double *qp, *pts, *dirs;
  double2 *cur, *part, *red;
};

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static size_t far_work_layout(const mom_mesh *m, int Q, const Group &G, int64_t ndir, char *base, FarWork *w) {
  const int64_t S = m->F * Q * G.n;
  const int64_t nchunk = (S + kFfChunk - 1) / kFfChunk;
  const size_t sizes[6] = {sizeof(double) * 3 * (size_t)m->F * Q, sizeof(double) * 3 * (size_t)S,
                           sizeof(double) * 8 * (size_t)ndir, sizeof(double2) * 6 * (size_t)S,
                           sizeof(double2) * 4 * (size_t)ndir * nchunk, sizeof(double2) * 4 * (size_t)ndir};
  size_t off[6], total = 0;
  for (int i = 0; i < 6; i++) {
    off[i] = total;
    total += align256(sizes[i]);
  }
  if (w) {
    w->qp = reinterpret_cast<double *>(base + off[0]);
    w->pts = reinterpret_cast<double *>(base + off[1]);
    w->dirs = reinterpret_cast<double *>(base + off[2]);
    w->cur = reinterpret_cast<double2 *>(base + off[3]);
    w->part = reinterpret_cast<double2 *>(base + off[4]);
    w->red = reinterpret_cast<double2 *>(base + off[5]);
  }
  return total;
}

size_t far_work_bytes(const mom_mesh *m, int Q, const Group &G, int64_t ndir) {
  return far_work_layout(m, Q, G, ndir, nullptr, nullptr);
}

// far_fields on the device: x the solution [2 Nt] (ldx per representation),
// theta, phi [ndir] and out [8 ndir] device arrays, work a device block of
// far_work_bytes bytes.  Every part of `work` is written before it is read,
// so it needs no clearing.  Launches only; the caller synchronises.
int far_fields_device(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, const double2 *x,
                      int64_t Nt, double r, int64_t ndir, const double *theta, const double *phi, double *out,
                      void *work, const Group &G, int64_t ldx) {
  const cd I(0.0, 1.0);
  const cd k = ri * omega / kC0;
  // the phasors exp(-i k dir . r) of every image point, |r| <= the mesh's
  // largest distance from the origin, reflected or not
  if (!(std::fabs(k.real()) * m->rmax < kCisMaxArg))
    return fail(MOM_ERR_ARG, "Re(k) times the mesh's distance from the origin is beyond the far field's phase range "
                             "(2^19 rad)");
  if (ndir < 1 || ndir > (1 << 24)) return fail(MOM_ERR_ARG, "number of directions out of range");
  FarWork w;
  far_work_layout(m, rule.Q, G, ndir, static_cast<char *>(work), &w);
  MOM_TRY(quad_points_into(ctx, m, rule, w.qp));
  const int64_t S = m->F * rule.Q * G.n;
  const int nchunk = (int)((S + kFfChunk - 1) / kFfChunk);
  const unsigned gd = (unsigned)((ndir + 255) / 256);
  hipLaunchKernelGGL(farfield_dirs_kernel, dim3(gd), dim3(256), 0, ctx->stream, ndir, theta, phi, w.dirs);
  hipLaunchKernelGGL(farfield_sources_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream, (int)m->F,
                     rule.Q, rule, w.qp, m->farea.as<double>(), sides(m, true), x, Nt, w.pts, w.cur, S, G,
                     ldx > 0 ? ldx : 2 * Nt);
  hipLaunchKernelGGL(farfield_dir_kernel, dim3(gd, (unsigned)nchunk), dim3(256), 0, ctx->stream, S, w.pts, w.cur,
                     (int)ndir, w.dirs, k.real(), k.imag(), w.part);
  hipLaunchKernelGGL(farfield_reduce_kernel, dim3(gd), dim3(256), 0, ctx->stream, (int)ndir, nchunk, w.part, w.red);
  FarCoef c;
  const cd P = I * k * std::exp(I * k * r) / (4.0 * MOM_PI * r), a = kEta0 / ri, b = ri / kEta0;
  c.P = make_double2(P.real(), P.imag());
  c.a = make_double2(a.real(), a.imag());
  c.b = make_double2(b.real(), b.imag());
  hipLaunchKernelGGL(farfield_combine_kernel, dim3(gd), dim3(256), 0, ctx->stream, ndir, w.red, c, out);
  MOM_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Transpose of the far-field map x -> out = (et, ep, ht, hp) of
// far_fields_device (identity group): for adjoint values v[8 d ..] of the
// outputs of direction d (complex, interleaved), g = L^T v, so that
// Re(g^T dx) = Re(sum_d v_d . dout_d).  The outputs fold back into the four
// radiation integrals per direction (farfield_adjoint_coef_kernel), the
// direction sum gives adjoint currents per point
// (farfield_adjoint_points_kernel), the currents project on each face's
// RWG sides (farfield_adjoint_faces_kernel), and every edge sums its two
// sides in face order (srcvec_gather_kernel).
//
// et = P (Lp + a Nt), ep = -P (Lt - a Np), ht = -P (Np - Lt b),
// hp = P (Nt + Lp b) give the adjoint integrals
// Nt' = P (a et' + hp'), Np' = P (a ep' - ht'), Lt' = P (b ht' - ep'),
// Lp' = P (et' + b hp').
#pragma clang fp contract(off)
__global__ void farfield_adjoint_coef_kernel(int64_t ndir, const double *__restrict__ v, FarCoef c,
                                             double2 *__restrict__ vb) {
  const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (d >= ndir) return;
  const double *o = v + 8 * d;
  const double2 et = make_double2(o[0], o[1]), ep = make_double2(o[2], o[3]), ht = make_double2(o[4], o[5]),
                hp = make_double2(o[6], o[7]);
  vb[4 * d] = zmul(c.P, zadd(zmul(c.a, et), hp));
  vb[4 * d + 1] = zmul(c.P, zsub(zmul(c.a, ep), ht));
  vb[4 * d + 2] = zmul(c.P, zsub(zmul(c.b, ht), ep));
  vb[4 * d + 3] = zmul(c.P, zadd(et, zmul(c.b, hp)));
}
#pragma clang fp contract(on)

constexpr int kFaTile = 128;  // directions per LDS tile of the adjoint point kernel

// One thread per point s: cbar[c][s] = sum_d exp(-i k dir_d . r_s)
// (u_theta Nt' + u_phi Np')_c (J, c < 3) and (u_theta Lt' + u_phi Lp')_c (M).
__global__ __launch_bounds__(256) void farfield_adjoint_points_kernel(int64_t S, const double *__restrict__ qp,
                                                                       int ndir, const double *__restrict__ dirs,
                                                                       const double2 *__restrict__ vb, double kr,
                                                                       double ki, double2 *__restrict__ cbar) {
  __shared__ double su[kFaTile][8];
  __shared__ double2 sv[kFaTile][4];
  const int tid = threadIdx.x;
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + tid;
  double r[3] = {0, 0, 0};
  if (s < S)
    for (int c = 0; c < 3; c++) r[c] = qp[3 * s + c];
  double J[3][2] = {{0, 0}, {0, 0}, {0, 0}}, M[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int d0 = 0; d0 < ndir; d0 += kFaTile) {
    const int nt = min(kFaTile, ndir - d0);
    if (tid < nt) {
      for (int c = 0; c < 8; c++) su[tid][c] = dirs[8 * (int64_t)(d0 + tid) + c];
      for (int k = 0; k < 4; k++) sv[tid][k] = vb[4 * (int64_t)(d0 + tid) + k];
    }
    __syncthreads();
    for (int q = 0; q < nt; q++) {
      const double *u = su[q];
      const double rd = u[0] * r[0] + u[1] * r[1] + u[2] * r[2];
      double er, ei;
      cis_exp(ki * rd, -kr * rd, er, ei);  // exp(-i k rd)
      double2 a[4];
      for (int k = 0; k < 4; k++)
        a[k] = make_double2(er * sv[q][k].x - ei * sv[q][k].y, er * sv[q][k].y + ei * sv[q][k].x);
      const double ut[3] = {u[3], u[4], u[5]}, up[3] = {u[6], u[7], 0.0};
      for (int c = 0; c < 3; c++) {
        J[c][0] += ut[c] * a[0].x + up[c] * a[1].x;
        J[c][1] += ut[c] * a[0].y + up[c] * a[1].y;
        M[c][0] += ut[c] * a[2].x + up[c] * a[3].x;
        M[c][1] += ut[c] * a[2].y + up[c] * a[3].y;
      }
    }
    __syncthreads();
  }
  if (s < S)
    for (int c = 0; c < 3; c++) {
      cbar[c * S + s] = make_double2(J[c][0], J[c][1]);
      cbar[(3 + c) * S + s] = make_double2(M[c][0], M[c][1]);
    }
}

// One thread per face: side q's adjoint J and M coefficients,
// w[2 (3f + q)] = sum_t w_t A f_q(r_t) . Jbar(r_t), w[2 (3f + q) + 1] likewise
// with Mbar (the transpose of farfield_sources_kernel).
__global__ void farfield_adjoint_faces_kernel(int F, int Q, QuadRule rule, const double *__restrict__ qp,
                                              const double *__restrict__ area, FaceSide side,
                                              const double2 *__restrict__ cbar, int64_t S, double2 *__restrict__ w) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  double aj[3][2] = {{0, 0}, {0, 0}, {0, 0}}, am[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int t = 0; t < Q; t++) {
    const int64_t s = (int64_t)f * Q + t;
    const double *r = qp + 3 * s;
    const double wa = rule.w[t] * area[f];
    double2 cj[3], cm[3];
    for (int c = 0; c < 3; c++) {
      cj[c] = cbar[c * S + s];
      cm[c] = cbar[(3 + c) * S + s];
    }
    for (int q = 0; q < 3; q++) {
      const double cf = side.coef[3 * f + q];
      const double *v = side.v + 3 * (3 * f + q);
      for (int c = 0; c < 3; c++) {
        const double fv = wa * (cf * (r[c] - v[c]));
        aj[q][0] += fv * cj[c].x;
        aj[q][1] += fv * cj[c].y;
        am[q][0] += fv * cm[c].x;
        am[q][1] += fv * cm[c].y;
      }
    }
  }
  for (int q = 0; q < 3; q++) {
    w[2 * (3 * f + q)] = make_double2(aj[q][0], aj[q][1]);
    w[2 * (3 * f + q) + 1] = make_double2(am[q][0], am[q][1]);
  }
}

// Per-side pairs w[3F][2] (J, M) summed per edge into the global vector
// g[2 Nt] (g[parent] and g[parent + Nt] += the edge's two sides, face order).
int gather_sides(mom_ctx *ctx, const mom_mesh *m, const double2 *w, int64_t Nt, double2 *g) {
  hipLaunchKernelGGL(srcvec_gather_kernel, dim3((unsigned)((m->E + 255) / 256)), dim3(256), 0, ctx->stream, m->E,
                     (int)m->F, (int64_t)1, m->eside.as<int>(), m->s_g.as<int>(), w, Nt, g, 2 * Nt);
  MOM_HIP(hipGetLastError());
  return 0;
}

int far_fields_adjoint(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t Nt,
                       double r, int64_t ndir, const double *theta, const double *phi, const double *v, double *g) {
  const cd I(0.0, 1.0);
  const cd k = ri * omega / kC0;
  if (!(std::fabs(k.real()) * m->rmax < kCisMaxArg))
    return fail(MOM_ERR_ARG, "Re(k) times the mesh's distance from the origin is beyond the far field's phase range "
                             "(2^19 rad)");
  if (ndir < 1 || ndir > (1 << 24)) return fail(MOM_ERR_ARG, "number of directions out of range");
  MOM_TRY(compute_quad_points(ctx, m, rule));
  const int64_t S = m->F * rule.Q;
  DevBuf dth, dph, ddirs, dv, dvb, dcb, dw, dg;
  int rc;
  if ((rc = dth.alloc(sizeof(double) * ndir)) || (rc = dph.alloc(sizeof(double) * ndir)) ||
      (rc = ddirs.alloc(sizeof(double) * 8 * ndir)) || (rc = dv.alloc(sizeof(double) * 8 * ndir)) ||
      (rc = dvb.alloc(sizeof(double2) * 4 * ndir)) || (rc = dcb.alloc(sizeof(double2) * 6 * S)) ||
      (rc = dw.alloc(sizeof(double2) * 6 * m->F)) || (rc = dg.alloc(sizeof(double2) * 2 * Nt)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dth.ptr, theta, sizeof(double) * ndir));
  MOM_TRY(copy_h2d(ctx, dph.ptr, phi, sizeof(double) * ndir));
  MOM_TRY(copy_h2d(ctx, dv.ptr, v, sizeof(double) * 8 * ndir));
  MOM_HIP(hipMemsetAsync(dg.ptr, 0, dg.bytes, ctx->stream));
  const unsigned gd = (unsigned)((ndir + 255) / 256);
  hipLaunchKernelGGL(farfield_dirs_kernel, dim3(gd), dim3(256), 0, ctx->stream, ndir, dth.as<double>(),
                     dph.as<double>(), ddirs.as<double>());
  FarCoef c;
  const cd P = I * k * std::exp(I * k * r) / (4.0 * MOM_PI * r), a = kEta0 / ri, b = ri / kEta0;
  c.P = make_double2(P.real(), P.imag());
  c.a = make_double2(a.real(), a.imag());
  c.b = make_double2(b.real(), b.imag());
  hipLaunchKernelGGL(farfield_adjoint_coef_kernel, dim3(gd), dim3(256), 0, ctx->stream, ndir, dv.as<double>(), c,
                     dvb.as<double2>());
  hipLaunchKernelGGL(farfield_adjoint_points_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream,
                     S, ctx->qp.as<double>(), (int)ndir, ddirs.as<double>(), dvb.as<double2>(), k.real(), k.imag(),
                     dcb.as<double2>());
  hipLaunchKernelGGL(farfield_adjoint_faces_kernel, dim3((unsigned)((m->F + 127) / 128)), dim3(128), 0, ctx->stream,
                     (int)m->F, rule.Q, rule, ctx->qp.as<double>(), m->farea.as<double>(), sides(m, true),
                     dcb.as<double2>(), S, dw.as<double2>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(gather_sides(ctx, m, dw.as<double2>(), Nt, dg.as<double2>()));
  return copy_d2h(ctx, g, dg.ptr, sizeof(double2) * 2 * Nt);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free far-field entry point: a fixed device array
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never uses the array's name as an address: its device address
// is looked up once per device at context creation (no_alloc_init) and
// kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

int no_alloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(MOM_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(MOM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return 0;
}

int far_fields_no_alloc(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                        const double2 *x, int64_t Nt, double r, int64_t ndir, const double *theta, const double *phi,
                        double *out) {
  void *base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  if (!base) return fail(MOM_ERR_STATE, "no-allocation scratch is not set up on this device");
  const Group G = identity_group();
  if (ndir < 1 || ndir > (1 << 24)) return fail(MOM_ERR_ARG, "number of directions out of range");
  const size_t need = far_work_bytes(m, rule.Q, G, ndir);
  if (need > kNoAllocScratchBytes)
    return fail(MOM_ERR_ARG, "far field of " + std::to_string(m->F) + " faces x " + std::to_string(rule.Q) +
                                 " points in " + std::to_string(ndir) + " directions needs " + std::to_string(need) +
                                 " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                                 " bytes of the no-allocation entry point");
  MOM_TRY(far_fields_device(ctx, m, rule, omega, ri, x, Nt, r, ndir, theta, phi, out, base, G, 0));
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Vertex-coordinate adjoint of the fill and of sigma_abs (no counterpart in
// the reference; the sensitivities of sysmat.f90:18-105, source.f90:136-247
// and cs.f90:14-104 with respect to the mesh nodes).
//
// With x the solution and lambda = Z^-T g (g = dsigma/dx), the derivative
// of sigma with respect to a node coordinate P is
//   dsigma/dP = dsigma/dP|_x + Re(lambda^T dq/dP) - Re(lambda^T (dZ/dP) x).
// Every term is a sum over faces or face pairs whose value depends only on
// the nodes of those faces.  The kernels below evaluate, per unordered face
// pair of one domain, Phi = sum_pq a_pq (D+H)_pq + b_pq K_pq over both
// orders of the pair (a = lambda_J x_J + Zsi lambda_M x_M, b = lambda_M x_J
// - lambda_J x_M at the pair's edges), rebuilt from the node positions with
// the fill's formulas (point pairs with G or the smoothed kernels, the
// analytic near parts, the self-pair rule), and differentiate it in each
// coordinate of each distinct node of the pair: the point-pair sums by
// reverse mode written out by hand (vfar_grad), the analytic near parts by
// forward mode (D1) through the int.f90 closed forms.  The check of both,
// mom_test_vertex_fd, takes the fourth-order central difference
// [8(f(h) - f(-h)) - (f(2h) - f(-2h))] / 12h with h = 1e-3 of the pair's
// mean edge length (truncation near 1e-12, rounding near 1e-10 relative).
// The near/far classification is frozen at the unperturbed mesh, as the
// reference's is piecewise constant in the nodes.
template <class T>
struct VFaceT {
  T P[3][3];        // nodes
  double sg[3];     // sign of each side's RWG function
  int lv[3];        // local index of each side's free vertex
};
using VFace = VFaceT<double>;

// node l's coordinate k by selects (a dynamic index would move the face to scratch)
template <class T>
__device__ __forceinline__ T vnode(const VFaceT<T> &f, int l, int k) {
  return l == 0 ? f.P[0][k] : (l == 1 ? f.P[1][k] : f.P[2][k]);
}

// One face's RWG data from its nodes (compute_basis_data, mesh.f90:1265-1322;
// rwgf.f90:34-78): area, coef = s L/(2A), div = s L/A, free vertex.
template <class T>
__device__ __forceinline__ void vface_rwg(const VFaceT<T> &f, T &A, T coef[3], T divv[3], T v[3][3]) {
  T a[3], b[3], c[3];
  for (int k = 0; k < 3; k++) {
    a[k] = f.P[2][k] - f.P[0][k];
    b[k] = f.P[2][k] - f.P[1][k];
  }
  cross3(a, b, c);
  A = 0.5 * norm3(c);
  for (int p = 0; p < 3; p++) {
    T e[3];
    for (int k = 0; k < 3; k++) e[k] = f.P[(p + 1) % 3][k] - f.P[p][k];
    const T L = norm3(e);
    coef[p] = f.sg[p] * (L / (2.0 * A));
    divv[p] = f.sg[p] * L / A;
    for (int k = 0; k < 3; k++) v[p][k] = vnode(f, f.lv[p], k);
  }
}

template <class T>
__device__ __forceinline__ void vface_point(const VFaceT<T> &f, const QuadRule &rule, int t, T r[3]) {
  const int Q = rule.Q;
  for (int k = 0; k < 3; k++)
    r[k] = rule.nodes[2 * Q + t] * f.P[0][k] + rule.nodes[t] * f.P[1][k] + rule.nodes[Q + t] * f.P[2][k];
}

template <class T>
__device__ void vface_near(const VFaceT<T> &f, NearFaceT<T> &o) {
  T a[3], b[3], c[3];
  for (int v = 0; v < 3; v++)
    for (int k = 0; k < 3; k++) o.p[v][k] = f.P[v][k];
  for (int k = 0; k < 3; k++) {
    a[k] = f.P[1][k] - f.P[0][k];
    b[k] = f.P[2][k] - f.P[0][k];
  }
  cross3(a, b, c);
  const T nc = norm3(c);
  for (int k = 0; k < 3; k++) o.n[k] = c[k] / nc;
  for (int e = 0; e < 3; e++) {
    T d[3];
    for (int k = 0; k < 3; k++) d[k] = f.P[(e + 1) % 3][k] - f.P[e][k];
    const T l = norm3(d);
    for (int k = 0; k < 3; k++) o.s[e][k] = d[k] / l;
    cross3(o.s[e], o.n, c);
    const T lm = norm3(c);
    for (int k = 0; k < 3; k++) o.m[e][k] = c[k] / lm;
  }
}

struct VPairArgs {
  int F;
  QuadRule rule;
  const double *p;     // nodes [V][3]
  const int *fnodes;   // [F][3]
  FaceSide side;       // global indices, coefficient signs, free vertices
  const double *cp;    // centroids (classification)
  double thr2;
  double kr, ki;
  double2 c1, c2, zsi, kk4, kk8;
  const double2 *x, *lam;  // [2N]
  int64_t N;
  int fd_only;         // differences for far pairs too (check of the closed form)
  const int *pairs;    // near pairs (test, source), sorted; the near launch takes those with test <= source
  const int64_t *near_off;  // [F + 1]: the near pairs of test face m are near_off[m] .. near_off[m+1]
  int64_t npairs;
  // Outputs, all by face: 9 values per face, Re d/d(coordinate c of the
  // face's local node v) at 3 v + c; every sum has a fixed order.
  double *part_t;      // [F][ntx][9]: pair tiles, test faces, summed over the tile's source faces
  double *part_s;      // [F][nty][9]: pair tiles, source faces, summed over the tile's test faces
  int ntx, nty;
  double *near_out;    // [npairs][18]: near pair (m <= n): m's 9 values, then n's
};

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Sum over the analytic near parts of test face ft against source face fs
// (near_fill_kernel's terms) contracted with a[3][3], b[3][3]: per test
// side p and source side q, a (c1 D + c2 H) + b K.  TT and TS, the scalar
// types of the test and the source face: double for the values, D1 in the
// face whose node coordinate the derivative is taken in.
template <class T>
struct CT {
  T x, y;
};
template <class T>
__device__ __forceinline__ CT<T> cmt(double2 a, CT<T> b) {
  return CT<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <class TT, class TS>
__device__ CT<PromoteT<TT, TS>> vnear_part_t(const VPairArgs &g, const VFaceT<TT> &ft, const VFaceT<TS> &fs,
                                             bool self, const double2 a[3][3], const double2 b[3][3]) {
  using T = PromoteT<TT, TS>;
  NearFaceT<TS> f;
  vface_near(fs, f);
  TT At, ct[3], dt[3], vt[3][3];
  TS As, cs[3], ds[3], vs[3][3];
  vface_rwg(ft, At, ct, dt, vt);
  vface_rwg(fs, As, cs, ds, vs);
  CT<T> acc{T(0.0), T(0.0)};
#pragma unroll 1
  for (int i = 0; i < g.rule.Q; i++) {
    TT r[3];
    vface_point(ft, g.rule, i, r);
    PointIntsT<T> P;
    d_point_ints(r, f, P);
    T k31[3], k3m1[3], rho[3];
    for (int c = 0; c < 3; c++) {
      T sm = f.m[0][c] * P.l1[0] + f.m[1][c] * P.l1[1] + f.m[2][c] * P.l1[2];
      k31[c] = (dval(P.h) == 0) ? sm : sm - P.h * f.n[c] * P.sm1;
      k3m1[c] = f.m[0][c] * P.lm1[0] + f.m[1][c] * P.lm1[1] + f.m[2][c] * P.lm1[2] + f.n[c] * P.sm3h;
      rho[c] = r[c] - P.h * f.n[c];
    }
    CT<T> s{T(0.0), T(0.0)};
    for (int q = 0; q < 3; q++) {
      const T h1 = ds[q] * P.sm1, h0 = ds[q] * P.s1;
      const CT<T> aH{g.kk4.x * h1 - g.kk8.x * h0, -(g.kk8.y * h0)};
      CT<T> aD[3], aK[3];
      for (int c = 0; c < 3; c++) {
        const T k2m1 = cs[q] * (f.m[0][c] * P.l1[0] + f.m[1][c] * P.l1[1] + f.m[2][c] * P.l1[2] +
                                (rho[c] - vs[q][c]) * P.sm1);
        const T k21 = cs[q] * (1.0 / 3.0 * (f.m[0][c] * P.l3[0] + f.m[1][c] * P.l3[1] + f.m[2][c] * P.l3[2]) +
                               (rho[c] - vs[q][c]) * P.s1);
        aD[c] = CT<T>{g.kk4.x * k2m1 - g.kk8.x * k21, -(g.kk8.y * k21)};
        aK[c] = CT<T>{T(0.0), T(0.0)};
      }
      if (!self) {
        T d[3] = {r[0] - vs[q][0], r[1] - vs[q][1], r[2] - vs[q][2]}, c1v[3], c0v[3];
        cross3(d, k3m1, c1v);
        cross3(d, k31, c0v);
        for (int c = 0; c < 3; c++) {
          const T k4m1 = -cs[q] * c1v[c], k41 = -cs[q] * c0v[c];
          aK[c] = CT<T>{g.kk4.x * k4m1 - g.kk8.x * k41, -(g.kk8.y * k41)};
        }
      }
      for (int p = 0; p < 3; p++) {
        TT fm[3];
        for (int c = 0; c < 3; c++) fm[c] = ct[p] * (r[c] - vt[p][c]);
        CT<T> dd{T(0.0), T(0.0)}, kk{T(0.0), T(0.0)};
        for (int c = 0; c < 3; c++) {
          dd.x += fm[c] * aD[c].x;
          dd.y += fm[c] * aD[c].y;
          kk.x += fm[c] * aK[c].x;
          kk.y += fm[c] * aK[c].y;
        }
        const CT<T> Dv = cmt(g.c1, dd);
        const CT<T> Hv = cmt(g.c2, CT<T>{dt[p] * aH.x, dt[p] * aH.y});
        const CT<T> t1 = cmt(a[p][q], CT<T>{Dv.x + Hv.x, Dv.y + Hv.y}), t2 = cmt(b[p][q], kk);
        s.x += t1.x + t2.x;
        s.y += t1.y + t2.y;
      }
    }
    acc.x += g.rule.w[i] * s.x * At;
    acc.y += g.rule.w[i] * s.y * At;
  }
  return acc;
}

__device__ double2 vnear_part(const VPairArgs &g, const VFace &ft, const VFace &fs, bool self, const double2 a[3][3],
                              const double2 b[3][3]) {
  const CT<double> v = vnear_part_t(g, ft, fs, self, a, b);
  return make_double2(v.x, v.y);
}

// The point-pair sums of the unordered pair (ft, fs) with the coefficients of
// both orders combined (a[p][q] for test side p of ft, source side q of fs).
__device__ double2 vfar_part(const VPairArgs &g, const VFace &ft, const VFace &fs, bool near, bool self,
                             const double2 a[3][3], const double2 b[3][3]) {
  double At, ct[3], dt[3], vt[3][3], As, cs[3], ds[3], vs[3][3];
  vface_rwg(ft, At, ct, dt, vt);
  vface_rwg(fs, As, cs, ds, vs);
  double2 T2 = make_double2(0.0, 0.0);
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      T2.x += a[p][q].x * dt[p] * ds[q];
      T2.y += a[p][q].y * dt[p] * ds[q];
    }
  const double2 c2T2 = cm(g.c2, T2);
  const double kr = g.kr, ki = g.ki, k2r = kr * kr - ki * ki, k2i = 2.0 * kr * ki;
  double2 acc = make_double2(0.0, 0.0);
  for (int i = 0; i < g.rule.Q; i++) {
    double r[3];
    vface_point(ft, g.rule, i, r);
    double u[3][3];
    for (int p = 0; p < 3; p++)
      for (int c = 0; c < 3; c++) u[p][c] = ct[p] * (r[c] - vt[p][c]);
    double2 X[3][3], Z[3][3];  // X_q = sum_p a_pq u_p, Z_q = sum_p b_pq u_p
    for (int q = 0; q < 3; q++)
      for (int c = 0; c < 3; c++) {
        X[q][c] = make_double2(0.0, 0.0);
        Z[q][c] = make_double2(0.0, 0.0);
        for (int p = 0; p < 3; p++) {
          X[q][c].x += a[p][q].x * u[p][c];
          X[q][c].y += a[p][q].y * u[p][c];
          Z[q][c].x += b[p][q].x * u[p][c];
          Z[q][c].y += b[p][q].y * u[p][c];
        }
      }
    double2 si = make_double2(0.0, 0.0);
    for (int j = 0; j < g.rule.Q; j++) {
      double rs[3];
      vface_point(fs, g.rule, j, rs);
      const double d[3] = {r[0] - rs[0], r[1] - rs[1], r[2] - rs[2]};
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double Gr, Gi, gr, gi;
      if (!near) {
        const double iR = rsqrt(R2), R = R2 * iR;
        double Er, Ei;
        cis_exp(-ki * R, kr * R, Er, Ei);
        const double e = iR * kOofpi;
        Gr = e * Er;
        Gi = e * Ei;
        const double ar = iR * iR + ki * iR, ai = -kr * iR;
        gr = Gr * ar - Gi * ai;
        gi = Gr * ai + Gi * ar;
      } else if (R2 == 0.0) {
        Gr = -ki * kOofpi;
        Gi = kr * kOofpi;
        gr = 0.0;
        gi = 0.0;
      } else {
        const double iR = rsqrt(R2), R = R2 * iR;
        double Er, Ei;
        cis_exp(-ki * R, kr * R, Er, Ei);
        Gr = ((Er - 1.0) * iR + 0.5 * k2r * R) * kOofpi;
        Gi = (Ei * iR + 0.5 * k2i * R) * kOofpi;
        const double iR2 = iR * iR, iR3 = iR2 * iR;
        const double ar = iR3 + ki * iR2, ai = -kr * iR2;
        gr = ((Er * ar - Ei * ai) - iR3 - 0.5 * k2r * iR) * kOofpi;
        gi = ((Er * ai + Ei * ar) - 0.5 * k2i * iR) * kOofpi;
      }
      double w[3][3];
      for (int q = 0; q < 3; q++)
        for (int c = 0; c < 3; c++) w[q][c] = cs[q] * (rs[c] - vs[q][c]);
      double2 T1 = make_double2(0.0, 0.0), T3 = make_double2(0.0, 0.0);
      for (int q = 0; q < 3; q++) {
        for (int c = 0; c < 3; c++) {
          T1.x += X[q][c].x * w[q][c];
          T1.y += X[q][c].y * w[q][c];
        }
        if (!self) {
          // d . (w_q x Z_q)
          for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            T3.x += d[c] * (w[q][c1] * Z[q][c2].x - w[q][c2] * Z[q][c1].x);
            T3.y += d[c] * (w[q][c1] * Z[q][c2].y - w[q][c2] * Z[q][c1].y);
          }
        }
      }
      const double2 G = make_double2(Gr, Gi), gg = make_double2(gr, gi);
      const double2 c1T1 = cm(g.c1, T1);
      const double2 v1 = cm(G, make_double2(c1T1.x + c2T2.x, c1T1.y + c2T2.y)), v2 = cm(gg, T3);
      si.x += g.rule.w[j] * (v1.x + v2.x);
      si.y += g.rule.w[j] * (v1.y + v2.y);
    }
    acc.x += g.rule.w[i] * si.x;
    acc.y += g.rule.w[i] * si.y;
  }
  const double AA = At * As;
  return make_double2(acc.x * AA, acc.y * AA);
}

__device__ void vload_face(const VPairArgs &g, int f, VFace &o) {
  for (int v = 0; v < 3; v++) {
    const int node = g.fnodes[3 * f + v];
    for (int k = 0; k < 3; k++) o.P[v][k] = g.p[3 * node + k];
  }
  for (int s = 0; s < 3; s++) {
    o.sg[s] = g.side.coef[3 * f + s] < 0 ? -1.0 : 1.0;
    const double *fv = g.side.v + 3 * (3 * f + s);
    int lv = 0;
    for (int v = 0; v < 3; v++)
      if (o.P[v][0] == fv[0] && o.P[v][1] == fv[1] && o.P[v][2] == fv[2]) lv = v;
    o.lv[s] = lv;
  }
}

// a, b of test face ft (sides p) and source face fs (sides q)
__device__ void vcoef(const VPairArgs &g, int ft, int fs, double2 a[3][3], double2 b[3][3]) {
  for (int p = 0; p < 3; p++) {
    const int64_t gi = g.side.g[3 * ft + p];
    const double2 lJ = g.lam[gi], lM = g.lam[gi + g.N];
    for (int q = 0; q < 3; q++) {
      const int64_t gj = g.side.g[3 * fs + q];
      const double2 xJ = g.x[gj], xM = g.x[gj + g.N];
      const double2 jj = cm(lJ, xJ), mm = cm(cm(g.zsi, lM), xM), mj = cm(lM, xJ), jm = cm(lJ, xM);
      a[p][q] = make_double2(jj.x + mm.x, jj.y + mm.y);
      b[p][q] = make_double2(mj.x - jm.x, mj.y - jm.y);
    }
  }
}

__device__ double2 vpair_value(const VPairArgs &g, const VFace &fm, const VFace &fn, bool near, bool self,
                               const double2 amn[3][3], const double2 bmn[3][3], const double2 anm[3][3],
                               const double2 bnm[3][3]) {
  double2 ac[3][3], bc[3][3];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      // (n, m) enters with its sides exchanged: (D+H) and K are symmetric in the faces
      ac[p][q] = self ? amn[p][q] : make_double2(amn[p][q].x + anm[q][p].x, amn[p][q].y + anm[q][p].y);
      bc[p][q] = self ? bmn[p][q] : make_double2(bmn[p][q].x + bnm[q][p].x, bmn[p][q].y + bnm[q][p].y);
    }
  double2 v = vfar_part(g, fm, fn, near, self, ac, bc);
  if (near) {
    const double2 n1 = vnear_part(g, fm, fn, self, amn, bmn);
    v.x += n1.x;
    v.y += n1.y;
    if (!self) {
      const double2 n2 = vnear_part(g, fn, fm, false, anm, bnm);
      v.x += n2.x;
      v.y += n2.y;
    }
  }
  return v;
}

// Far pairs: the derivative of the point-pair sum in closed form (reverse
// mode by hand).  With e_p = r - v_p, f_q = r' - v_q, d = r - r' (r, r' the
// test and source points, v the free vertices), alpha = s L/2 and beta =
// s L (the areas cancel against the RWG coefficients), each point pair adds
//   G (c1 sum_pq a_pq alpha_p alpha_q e_p.f_q + c2 sum_pq a_pq beta_p beta_q)
//   + g sum_pq b_pq alpha_p alpha_q d.(f_q x e_p).
// With A_pq = a_pq alpha_p alpha_q and B_pq = b_pq alpha_p alpha_q the sums
// are polynomials in r and r' with per-pair coefficients,
//   sum A e_p.f_q = S r.r' - r.Vs - Vt.r' + W,
//   sum B d.(f_q x e_p) = Y.(r x r') + d.X,
// S = sum A, Vt = sum A v_p, Vs = sum A v_q, W = sum A v_p.v_q,
// Y = sum B (v_q - v_p), X = sum B (v_q x v_p).  The point loop therefore
// carries only the gradients in r and r' (to the barycentric points) and 14
// complex moments of G and g, which give the derivatives in the
// coefficients; those go to the edge lengths and free vertices per pair.
// G' = G (ik - 1/R) and g' = G (k^2/R - 3/R^3 + 3ik/R^2).  gt/gs: Re
// d/d(node) of the test and source faces' nodes (local order).
// NEAR: the smoothed kernels of near pairs, Gs = [(e^{ikR} - 1)/R + k^2 R/2]/4pi
// and gs = [e^{ikR}(1/R^3 - ik/R^2) - 1/R^3 - k^2/(2R)]/4pi, with
// Gs' = [ik e^{ikR}/R - (e^{ikR} - 1)/R^2 + k^2/2]/4pi and
// gs' = [ik e^{ikR}(1/R^3 - ik/R^2) + e^{ikR}(2ik/R^3 - 3/R^4) + 3/R^4 + k^2/(2R^2)]/4pi;
// coincident points (R = 0, self pairs) contribute Gs(0) = ik/4pi and no
// derivative.  A self pair passes b = 0 (its K block is zero).
__device__ __forceinline__ double re_mul(double2 a, double2 b) { return a.x * b.x - a.y * b.y; }

// The coefficients of the unordered pair (m, n) with both orders combined
// (a[p][q] for side p of m and side q of n); a self pair takes its own a and
// b = 0 (its K block is zero).
__device__ void vcoef_pair(const VPairArgs &g, int m, int n, double2 a[3][3], double2 b[3][3]) {
  vcoef(g, m, n, a, b);
  if (m == n) {
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) b[p][q] = make_double2(0.0, 0.0);
    return;
  }
  double2 an[3][3], bn[3][3];
  vcoef(g, n, m, an, bn);
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      a[p][q] = make_double2(a[p][q].x + an[q][p].x, a[p][q].y + an[q][p].y);
      b[p][q] = make_double2(b[p][q].x + bn[q][p].x, b[p][q].y + bn[q][p].y);
    }
}

template <bool NEAR>
__device__ void vfar_grad(const VPairArgs &g, int m, int n, const VFace &ft, const VFace &fs, double gt[3][3],
                          double gs[3][3]) {
  double al_t[3], al_s[3];
  for (int p = 0; p < 3; p++) {
    double et[3], es[3];
    for (int k = 0; k < 3; k++) {
      et[k] = ft.P[(p + 1) % 3][k] - ft.P[p][k];
      es[k] = fs.P[(p + 1) % 3][k] - fs.P[p][k];
    }
    al_t[p] = ft.sg[p] * norm3(et) / 2.0;
    al_s[p] = fs.sg[p] * norm3(es) / 2.0;
  }
  // the pair's coefficients, with c1 and c2 folded in where the loop uses them
  double2 c1S, c1Vs[3], c1Vt[3], K0, Y[3], X[3];
  {
    double2 a[3][3], b[3][3];
    vcoef_pair(g, m, n, a, b);
    double2 S = make_double2(0.0, 0.0), W = S, Vs[3], Vt[3];
    for (int c = 0; c < 3; c++) Vs[c] = Vt[c] = Y[c] = X[c] = S;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) {
        const double aa = al_t[p] * al_s[q];
        const double2 At = make_double2(a[p][q].x * aa, a[p][q].y * aa);
        const double2 Bt = make_double2(b[p][q].x * aa, b[p][q].y * aa);
        double vp[3], vq[3];
        for (int c = 0; c < 3; c++) {
          vp[c] = vnode(ft, ft.lv[p], c);
          vq[c] = vnode(fs, fs.lv[q], c);
        }
        double x[3];
        cross3(vq, vp, x);
        const double vv = vp[0] * vq[0] + vp[1] * vq[1] + vp[2] * vq[2];
        S.x += At.x;
        S.y += At.y;
        W.x += At.x * vv;
        W.y += At.y * vv;
        for (int c = 0; c < 3; c++) {
          Vt[c].x += At.x * vp[c];
          Vt[c].y += At.y * vp[c];
          Vs[c].x += At.x * vq[c];
          Vs[c].y += At.y * vq[c];
          Y[c].x += Bt.x * (vq[c] - vp[c]);
          Y[c].y += Bt.y * (vq[c] - vp[c]);
          X[c].x += Bt.x * x[c];
          X[c].y += Bt.y * x[c];
        }
      }
    c1S = cm(g.c1, S);
    for (int c = 0; c < 3; c++) {
      c1Vs[c] = cm(g.c1, Vs[c]);
      c1Vt[c] = cm(g.c1, Vt[c]);
    }
    const double2 c1W = cm(g.c1, W), c2S = cm(g.c2, S);
    K0 = make_double2(c1W.x + 4.0 * c2S.x, c1W.y + 4.0 * c2S.y);
  }
  const double kr = g.kr, ki = g.ki, k2r = kr * kr - ki * ki, k2i = 2.0 * kr * ki;
  // GT/GS: Re d/d(node) through the points; moments of G and g
  double GT[3][3], GS[3][3];
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) GT[k][c] = GS[k][c] = 0.0;
  double2 MG = make_double2(0.0, 0.0), Mrr = MG, MGr[3], MGs[3], Mgx[3], Mgd[3];
  for (int c = 0; c < 3; c++) MGr[c] = MGs[c] = Mgx[c] = Mgd[c] = MG;
  const int Q = g.rule.Q;
  for (int i = 0; i < Q; i++) {
    double r[3];
    vface_point(ft, g.rule, i, r);
    double dRi[3] = {0.0, 0.0, 0.0};
    double2 sG = make_double2(0.0, 0.0), sg = sG, sGr[3], sgr[3];
    for (int c = 0; c < 3; c++) sGr[c] = sgr[c] = sG;
    for (int j = 0; j < Q; j++) {
      double rs[3];
      vface_point(fs, g.rule, j, rs);
      const double d[3] = {r[0] - rs[0], r[1] - rs[1], r[2] - rs[2]};
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double iR;
      double2 G, gg, Gp, gp;
      if (!NEAR) {
        iR = rsqrt(R2);
        const double R = R2 * iR;
        double Er, Ei;
        cis_exp(-ki * R, kr * R, Er, Ei);
        G = make_double2(iR * kOofpi * Er, iR * kOofpi * Ei);
        gg = cm(G, make_double2(iR * iR + ki * iR, -kr * iR));
        // G' = G (ik - 1/R) = G (-ki - 1/R + i kr)
        Gp = cm(G, make_double2(-ki - iR, kr));
        // g' = G (k^2/R - 3/R^3 + 3ik/R^2)
        gp = cm(G, make_double2(k2r * iR - 3.0 * iR * iR * iR - 3.0 * ki * iR * iR, k2i * iR + 3.0 * kr * iR * iR));
      } else if (R2 == 0.0) {
        iR = 0.0;
        G = make_double2(-ki * kOofpi, kr * kOofpi);
        gg = Gp = gp = make_double2(0.0, 0.0);
      } else {
        iR = rsqrt(R2);
        const double R = R2 * iR, iR2 = iR * iR, iR3 = iR2 * iR, iR4 = iR2 * iR2;
        double Er, Ei;
        cis_exp(-ki * R, kr * R, Er, Ei);
        const double2 E = make_double2(Er, Ei), ik = make_double2(-ki, kr);
        const double2 Em1 = make_double2(Er - 1.0, Ei);
        G = make_double2((Em1.x * iR + 0.5 * k2r * R) * kOofpi, (Em1.y * iR + 0.5 * k2i * R) * kOofpi);
        const double2 h = make_double2(iR3 + ki * iR2, -kr * iR2);  // 1/R^3 - ik/R^2
        const double2 Eh = cm(E, h);
        gg = make_double2((Eh.x - iR3 - 0.5 * k2r * iR) * kOofpi, (Eh.y - 0.5 * k2i * iR) * kOofpi);
        const double2 ikE = cm(ik, E);
        Gp = make_double2((ikE.x * iR - Em1.x * iR2 + 0.5 * k2r) * kOofpi, (ikE.y * iR - Em1.y * iR2 + 0.5 * k2i) * kOofpi);
        const double2 ikEh = cm(ikE, h);
        const double2 Ehp = cm(E, make_double2(-3.0 * iR4 - 2.0 * ki * iR3, 2.0 * kr * iR3));  // e^{ikR}(2ik/R^3 - 3/R^4)
        gp = make_double2((ikEh.x + Ehp.x + 3.0 * iR4 + 0.5 * k2r * iR2) * kOofpi,
                          (ikEh.y + Ehp.y + 0.5 * k2i * iR2) * kOofpi);
      }
      // P1 = c1 (S r.r' - r.Vs - Vt.r' + W) + 4 c2 S, P2 = Y.(r x r') + d.X
      double rxs[3];
      cross3(r, rs, rxs);
      const double rr = r[0] * rs[0] + r[1] * rs[1] + r[2] * rs[2];
      double2 P1 = make_double2(c1S.x * rr + K0.x, c1S.y * rr + K0.y), P2 = make_double2(0.0, 0.0);
      for (int c = 0; c < 3; c++) {
        P1.x -= r[c] * c1Vs[c].x + rs[c] * c1Vt[c].x;
        P1.y -= r[c] * c1Vs[c].y + rs[c] * c1Vt[c].y;
        P2.x += Y[c].x * rxs[c] + d[c] * X[c].x;
        P2.y += Y[c].y * rxs[c] + d[c] * X[c].y;
      }
      // Re dPsi/dR / R, and the parts of Re dPsi/dr, dPsi/dr'
      const double hR = (re_mul(Gp, P1) + re_mul(gp, P2)) * iR;
      const double GSr = re_mul(G, c1S);
      double gY[3], Dr[3], Ds[3];
      for (int c = 0; c < 3; c++) gY[c] = re_mul(gg, Y[c]);
      double yxr[3], sxy[3];
      cross3(rs, gY, sxy);
      cross3(gY, r, yxr);
      const double ww = g.rule.w[i] * g.rule.w[j], wj = g.rule.w[j];
      for (int c = 0; c < 3; c++) {
        const double gX = re_mul(gg, X[c]);
        Dr[c] = hR * d[c] + GSr * rs[c] - re_mul(G, c1Vs[c]) + sxy[c] + gX;
        Ds[c] = -hR * d[c] + GSr * r[c] - re_mul(G, c1Vt[c]) + yxr[c] - gX;
        dRi[c] += wj * Dr[c];
      }
      const double bj[3] = {g.rule.nodes[2 * Q + j], g.rule.nodes[j], g.rule.nodes[Q + j]};
      for (int k = 0; k < 3; k++) {
        const double wb = ww * bj[k];
        for (int c = 0; c < 3; c++) GS[k][c] += wb * Ds[c];
      }
      sG.x += wj * G.x;
      sG.y += wj * G.y;
      sg.x += wj * gg.x;
      sg.y += wj * gg.y;
      for (int c = 0; c < 3; c++) {
        sGr[c].x += wj * G.x * rs[c];
        sGr[c].y += wj * G.y * rs[c];
        sgr[c].x += wj * gg.x * rs[c];
        sgr[c].y += wj * gg.y * rs[c];
      }
    }
    const double wi = g.rule.w[i];
    const double bi[3] = {g.rule.nodes[2 * Q + i], g.rule.nodes[i], g.rule.nodes[Q + i]};
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) GT[k][c] += wi * bi[k] * dRi[c];
    MG.x += wi * sG.x;
    MG.y += wi * sG.y;
    for (int c = 0; c < 3; c++) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      MGr[c].x += wi * r[c] * sG.x;
      MGr[c].y += wi * r[c] * sG.y;
      MGs[c].x += wi * sGr[c].x;
      MGs[c].y += wi * sGr[c].y;
      Mrr.x += wi * r[c] * sGr[c].x;
      Mrr.y += wi * r[c] * sGr[c].y;
      Mgx[c].x += wi * (r[c1] * sgr[c2].x - r[c2] * sgr[c1].x);
      Mgx[c].y += wi * (r[c1] * sgr[c2].y - r[c2] * sgr[c1].y);
      Mgd[c].x += wi * (r[c] * sg.x - sgr[c].x);
      Mgd[c].y += wi * (r[c] * sg.y - sgr[c].y);
    }
  }
  // derivatives in the coefficients: dS = c1 Mrr + 4 c2 MG, dVs = -c1 MGr,
  // dVt = -c1 MGs, dW = c1 MG, dY = Mgx, dX = Mgd
  double2 DS, DVs[3], DVt[3], DW;
  {
    const double2 a1 = cm(g.c1, Mrr), a2 = cm(g.c2, MG);
    DS = make_double2(a1.x + 4.0 * a2.x, a1.y + 4.0 * a2.y);
    DW = cm(g.c1, MG);
    for (int c = 0; c < 3; c++) {
      const double2 t1 = cm(g.c1, MGr[c]), t2 = cm(g.c1, MGs[c]);
      DVs[c] = make_double2(-t1.x, -t1.y);
      DVt[c] = make_double2(-t2.x, -t2.y);
    }
  }
  // ... and through them in the edge lengths (dAt, dAs) and free vertices;
  // the a, b are read again rather than held through the point loop (the
  // empty asm keeps the compiler from reusing the first reads)
  asm volatile("" ::: "memory");
  double2 a[3][3], b[3][3];
  vcoef_pair(g, m, n, a, b);
  double dAt[3] = {0.0, 0.0, 0.0}, dAs[3] = {0.0, 0.0, 0.0}, VT[3][3], VS[3][3];
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) VT[k][c] = VS[k][c] = 0.0;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      double vp[3], vq[3];
      for (int c = 0; c < 3; c++) {
        vp[c] = vnode(ft, ft.lv[p], c);
        vq[c] = vnode(fs, fs.lv[q], c);
      }
      double x[3];
      cross3(vq, vp, x);
      const double vv = vp[0] * vq[0] + vp[1] * vq[1] + vp[2] * vq[2];
      // Om = a (DS + DVt.v_p + DVs.v_q + DW v_p.v_q) + b (DY.(v_q - v_p) + DX.(v_q x v_p))
      double2 oa = make_double2(DS.x + DW.x * vv, DS.y + DW.y * vv), ob = make_double2(0.0, 0.0);
      for (int c = 0; c < 3; c++) {
        oa.x += DVt[c].x * vp[c] + DVs[c].x * vq[c];
        oa.y += DVt[c].y * vp[c] + DVs[c].y * vq[c];
        ob.x += Mgx[c].x * (vq[c] - vp[c]) + Mgd[c].x * x[c];
        ob.y += Mgx[c].y * (vq[c] - vp[c]) + Mgd[c].y * x[c];
      }
      const double om = re_mul(a[p][q], oa) + re_mul(b[p][q], ob);
      dAt[p] += al_s[q] * om;
      dAs[q] += al_t[p] * om;
      // d/dv_p: a (DVt + DW v_q) + b (-DY + DX x v_q); d/dv_q: a (DVs + DW v_p) + b (DY + v_p x DX)
      const double aa = al_t[p] * al_s[q];
      for (int c = 0; c < 3; c++) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const double2 xq = make_double2(Mgd[c1].x * vq[c2] - Mgd[c2].x * vq[c1], Mgd[c1].y * vq[c2] - Mgd[c2].y * vq[c1]);
        const double2 px = make_double2(vp[c1] * Mgd[c2].x - vp[c2] * Mgd[c1].x, vp[c1] * Mgd[c2].y - vp[c2] * Mgd[c1].y);
        const double tp = re_mul(a[p][q], make_double2(DVt[c].x + DW.x * vq[c], DVt[c].y + DW.y * vq[c])) +
                          re_mul(b[p][q], make_double2(xq.x - Mgx[c].x, xq.y - Mgx[c].y));
        const double tq = re_mul(a[p][q], make_double2(DVs[c].x + DW.x * vp[c], DVs[c].y + DW.y * vp[c])) +
                          re_mul(b[p][q], make_double2(Mgx[c].x + px.x, Mgx[c].y + px.y));
        VT[p][c] += aa * tp;
        VS[q][c] += aa * tq;
      }
    }
  // alpha_p = s_p L_p / 2 with L_p = |P_{p+1} - P_p|
  for (int p = 0; p < 3; p++) {
    double et[3], es[3];
    for (int k = 0; k < 3; k++) {
      et[k] = ft.P[(p + 1) % 3][k] - ft.P[p][k];
      es[k] = fs.P[(p + 1) % 3][k] - fs.P[p][k];
    }
    const double lt = norm3(et), ls = norm3(es);
    for (int c = 0; c < 3; c++) {
      const double ht = ft.sg[p] * 0.5 * et[c] / lt, hs = fs.sg[p] * 0.5 * es[c] / ls;
      GT[(p + 1) % 3][c] += dAt[p] * ht;
      GT[p][c] -= dAt[p] * ht;
      GS[(p + 1) % 3][c] += dAs[p] * hs;
      GS[p][c] -= dAs[p] * hs;
    }
  }
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) {
      double t = GT[k][c], u = GS[k][c];
      for (int p = 0; p < 3; p++) {
        t += ft.lv[p] == k ? VT[p][c] : 0.0;
        u += fs.lv[p] == k ? VS[p][c] : 0.0;
      }
      gt[k][c] = t;
      gs[k][c] = u;
    }
}

// Re d Phi / d(node coordinate) of the unordered face pair (m <= n) by
// fourth-order central differences in the coordinates of its distinct
// nodes, added to ot[3 v + c] (face m's local node v) and os (face n's; a
// node of both counts in ot, and a self pair adds only to ot).  The slots
// are this thread's own (LDS or global memory), so nothing stays in
// registers across the differences.  This is the check of the closed forms
// (mom_test_vertex_fd).
__device__ void vertex_pair(const VPairArgs &g, int m, int n, bool near, double *ot, double *os) {
  const bool self = (m == n);
  VFace fm, fn;
  vload_face(g, m, fm);
  vload_face(g, n, fn);
  double2 amn[3][3], bmn[3][3], anm[3][3], bnm[3][3];
  vcoef(g, m, n, amn, bmn);
  vcoef(g, n, m, anm, bnm);
  // distinct nodes and their places in the two faces
  int nodes[6], nn = 0;
  for (int v = 0; v < 6; v++) {
    const int node = v < 3 ? g.fnodes[3 * m + v] : g.fnodes[3 * n + v - 3];
    bool seen = false;
    for (int k = 0; k < nn; k++) seen |= nodes[k] == node;
    if (!seen) nodes[nn++] = node;
  }
  double L = 0;
  for (int v = 0; v < 3; v++) {
    double e1[3], e2[3];
    for (int k = 0; k < 3; k++) {
      e1[k] = fm.P[(v + 1) % 3][k] - fm.P[v][k];
      e2[k] = fn.P[(v + 1) % 3][k] - fn.P[v][k];
    }
    L += norm3(e1) + norm3(e2);
  }
  const double h = 1e-3 * L / 6.0;
  const double steps[4] = {h, -h, 2.0 * h, -2.0 * h};
  for (int k = 0; k < nn; k++) {
    const int node = nodes[k];
    for (int c = 0; c < 3; c++) {
      double val[4];
      for (int st = 0; st < 4; st++) {
        VFace pm = fm, pn = fn;
        for (int v = 0; v < 3; v++) {
          if (g.fnodes[3 * m + v] == node) pm.P[v][c] += steps[st];
          if (g.fnodes[3 * n + v] == node) pn.P[v][c] += steps[st];
        }
        val[st] = vpair_value(g, pm, pn, near, self, amn, bmn, anm, bnm).x;
      }
      const double dv = (8.0 * (val[0] - val[1]) - (val[2] - val[3])) / (12.0 * h);
      int vm = -1, vn = -1;
      for (int v = 0; v < 3; v++) {
        if (g.fnodes[3 * m + v] == node) vm = v;
        if (g.fnodes[3 * n + v] == node) vn = v;
      }
      if (vm >= 0)
        ot[3 * vm + c] += dv;
      else if (vn >= 0)
        os[3 * vn + c] += dv;
    }
  }
}

// Pair tiles of kVTM test faces x kVTN source faces (one thread per pair,
// the pairs m <= n).  FD = false: the far pairs in closed form; the near
// pairs have their own launch, so that their differences do not hold whole
// waves of far pairs.  FD = true: every pair by differences (the check of
// the closed form).  Each tile sums its test faces' values over its source
// faces and its source faces' values over its test faces, in a fixed order,
// into part_t / part_s (no atomics: the gradient is the same bit for bit
// from run to run).
constexpr int kVTM = 8, kVTN = 16;
template <bool FD>
__global__ __launch_bounds__(128) void vertex_pair_kernel(VPairArgs g) {
  __shared__ double st[kVTM][kVTN][9], ss[kVTM][kVTN][9];
  const int tn = threadIdx.x % kVTN, tm = threadIdx.x / kVTN;
  const int n0 = blockIdx.x * kVTN, m0 = blockIdx.y * kVTM;
  if (n0 + kVTN - 1 < m0) return;  // the whole tile below the diagonal (uniform)
  const int m = m0 + tm, n = n0 + tn;
  for (int k = 0; k < 9; k++) st[tm][tn][k] = ss[tm][tn][k] = 0.0;
  if (m < g.F && n < g.F && n >= m) {
    const bool near = near_pair(g.cp + 3 * m, g.cp + 3 * n, g.thr2);
    if (FD) {
      vertex_pair(g, m, n, near, &st[tm][tn][0], &ss[tm][tn][0]);
    } else if (!near) {
      VFace fm, fn;
      vload_face(g, m, fm);
      vload_face(g, n, fn);
      double gt[3][3], gs[3][3];
      vfar_grad<false>(g, m, n, fm, fn, gt, gs);
      for (int v = 0; v < 3; v++)
        for (int c = 0; c < 3; c++) {
          st[tm][tn][3 * v + c] = gt[v][c];
          ss[tm][tn][3 * v + c] = gs[v][c];
        }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kVTM * 9; e += 128) {
    const int r = e / 9, k = e % 9;
    if (m0 + r >= g.F) continue;
    double t = 0.0;
    for (int j = 0; j < kVTN; j++) t += st[r][j][k];
    g.part_t[((int64_t)(m0 + r) * g.ntx + blockIdx.x) * 9 + k] = t;
  }
  for (int e = threadIdx.x; e < kVTN * 9; e += 128) {
    const int j = e / 9, k = e % 9;
    if (n0 + j >= g.F) continue;
    double t = 0.0;
    for (int r = 0; r < kVTM; r++) t += ss[r][j][k];
    g.part_s[((int64_t)(n0 + j) * g.nty + blockIdx.y) * 9 + k] = t;
  }
}

// One thread per entry of the near list with test <= source: the smoothed
// point loop in closed form; the pair's values go to its entry of near_out
// (vertex_near_analytic_kernel adds the int.f90 analytic parts after).
__global__ __launch_bounds__(128) void vertex_near_kernel(VPairArgs g) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= g.npairs) return;
  const int m = g.pairs[2 * id], n = g.pairs[2 * id + 1];
  if (n < m) return;
  const bool self = (m == n);
  {
    VFace fm, fn;
    vload_face(g, m, fm);
    vload_face(g, n, fn);
    double gt[3][3], gs[3][3];
    vfar_grad<true>(g, m, n, fm, fn, gt, gs);
    if (self)  // both copies of the face are face m
      for (int v = 0; v < 3; v++)
        for (int c = 0; c < 3; c++) {
          gt[v][c] += gs[v][c];
          gs[v][c] = 0.0;
        }
    double *o = g.near_out + 18 * id;
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) {
        o[3 * v + c] = gt[v][c];
        o[9 + 3 * v + c] = gs[v][c];
      }
  }
}

// The int.f90 analytic parts of the near pairs, differentiated exactly:
// forward mode (D1) through the faces' geometry, the closed forms and the
// contraction with the coefficients, one thread per (entry m <= n of the
// near list, local vertex v of a face, coordinate c).  Three launches, each
// with one code path for all its threads: ROLE 0 the nodes of face m that
// face n does not hold (face n stays in plain doubles), ROLE 1 those of
// face n that face m does not hold, ROLE 2 the nodes of face m that face n
// holds too (every node of a self pair; both faces in D1).  The derivative
// goes to the thread's own slot of the entry (face m's local node v for
// ROLE 0 and 2, face n's for ROLE 1), added to what vertex_near_kernel
// wrote there.
__device__ __forceinline__ void vseed(const VFace &f0, int node, int c, const int *fnodes, VFaceT<D1> &f) {
  for (int v = 0; v < 3; v++) {
    const bool on = fnodes[v] == node;
    for (int j = 0; j < 3; j++) f.P[v][j] = D1(f0.P[v][j], (on && j == c) ? 1.0 : 0.0);
    f.sg[v] = f0.sg[v];
    f.lv[v] = f0.lv[v];
  }
}

template <int ROLE>
__global__ __launch_bounds__(128) void vertex_near_analytic_kernel(VPairArgs g) {
  // the coefficients live in LDS (each thread its own), not in registers
  __shared__ double2 sab[128][2][3][3];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t id = t / 9;
  if (id >= g.npairs) return;
  const int sub = (int)(t - 9 * id), v = sub / 3, c = sub - 3 * v;
  const int m = g.pairs[2 * id], n = g.pairs[2 * id + 1];
  if (n < m) return;
  const int own = ROLE == 1 ? n : m, other = ROLE == 1 ? m : n;
  const int node = g.fnodes[3 * own + v];
  const bool shared = g.fnodes[3 * other] == node || g.fnodes[3 * other + 1] == node || g.fnodes[3 * other + 2] == node;
  if (shared != (ROLE == 2)) return;
  VFace fm0, fn0;
  vload_face(g, m, fm0);
  vload_face(g, n, fn0);
  double2(*a)[3] = sab[threadIdx.x][0], (*b)[3] = sab[threadIdx.x][1];
  double dv;
  if (ROLE == 2) {
    const bool self = (m == n);
    VFaceT<D1> fm, fn;
    vseed(fm0, node, c, g.fnodes + 3 * m, fm);
    vseed(fn0, node, c, g.fnodes + 3 * n, fn);
    vcoef(g, m, n, a, b);
    dv = vnear_part_t(g, fm, fn, self, a, b).x.d;
    if (!self) {
      vcoef(g, n, m, a, b);
      dv += vnear_part_t(g, fn, fm, false, a, b).x.d;
    }
  } else if (ROLE == 0) {
    VFaceT<D1> fm;
    vseed(fm0, node, c, g.fnodes + 3 * m, fm);
    vcoef(g, m, n, a, b);
    dv = vnear_part_t(g, fm, fn0, false, a, b).x.d;
    vcoef(g, n, m, a, b);
    dv += vnear_part_t(g, fn0, fm, false, a, b).x.d;
  } else {
    VFaceT<D1> fn;
    vseed(fn0, node, c, g.fnodes + 3 * n, fn);
    vcoef(g, m, n, a, b);
    dv = vnear_part_t(g, fm0, fn, false, a, b).x.d;
    vcoef(g, n, m, a, b);
    dv += vnear_part_t(g, fn, fm0, false, a, b).x.d;
  }
  g.near_out[18 * id + (ROLE == 1 ? 9 : 0) + 3 * v + c] += dv;
}

// Per face: its tile partial sums (in tile order) and its near pairs (in
// near-list order; a pair listed as (f, n) with n < f has its values in the
// entry (n, f), found by bisection in n's sorted row), into fg[F][9].
__global__ __launch_bounds__(128) void vertex_face_sum_kernel(VPairArgs g, double *fg) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.F) return;
  double acc[9];
  for (int k = 0; k < 9; k++) acc[k] = 0.0;
  for (int t = 0; t < g.ntx; t++)
    for (int k = 0; k < 9; k++) acc[k] += g.part_t[((int64_t)f * g.ntx + t) * 9 + k];
  for (int t = 0; t < g.nty; t++)
    for (int k = 0; k < 9; k++) acc[k] += g.part_s[((int64_t)f * g.nty + t) * 9 + k];
  if (g.near_out) {
    for (int64_t id = g.near_off[f]; id < g.near_off[f + 1]; id++) {
      const int n = g.pairs[2 * id + 1];
      const double *o = nullptr;
      if (n >= f) {
        o = g.near_out + 18 * id;
      } else {
        int64_t lo = g.near_off[n], hi = g.near_off[n + 1] - 1;
        while (lo < hi) {
          const int64_t mid = (lo + hi) / 2;
          if (g.pairs[2 * mid + 1] < f)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo < g.near_off[n + 1] && g.pairs[2 * lo + 1] == f) o = g.near_out + 18 * lo + 9;
      }
      if (o)
        for (int k = 0; k < 9; k++) acc[k] += o[k];
    }
  }
  for (int k = 0; k < 9; k++) fg[(int64_t)f * 9 + k] = acc[k];
}

// Per face of the exterior domain: Psi = the face's share of sigma_abs
// (cs_prtsrf's int1 term, cs.f90:38-102) + Re sum_p lambda_J q^E_p +
// lambda_M q^H_p (srcvec's terms, source.f90:224-244), differentiated like
// the pairs.
struct VFaceArgs {
  int F;
  QuadRule rule;
  const double *p;
  const int *fnodes;
  FaceSide side;
  PwSrc S;
  double kr, ki;
  double2 hfac;
  double cabs_c;  // -2/(eps0 c0 Re ri)
  const double2 *x, *lam;
  int64_t N;
  double *fg;  // [F][9]: Re d/d(coordinate c of the face's local node v) at 3 v + c
};

__device__ double vface_value(const VFaceArgs &g, int f, const VFace &F) {
  double A, coef[3], divv[3], v[3][3];
  vface_rwg(F, A, coef, divv, v);
  double a[3], b[3], nr[3];
  for (int k = 0; k < 3; k++) {
    a[k] = F.P[1][k] - F.P[0][k];
    b[k] = F.P[2][k] - F.P[0][k];
  }
  cross3(a, b, nr);
  const double nl = norm3(nr);
  for (int k = 0; k < 3; k++) nr[k] /= nl;
  double int1 = 0.0;
  double2 src = make_double2(0.0, 0.0);
  for (int t = 0; t < g.rule.Q; t++) {
    double r[3];
    vface_point(F, g.rule, t, r);
    double Et[3][2] = {{0, 0}, {0, 0}, {0, 0}}, Ht[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    double Ei[3][2], Hi[3][2];
    pw_eh(g.S, r, g.kr, g.ki, g.hfac, Ei, Hi);
    for (int p = 0; p < 3; p++) {
      double fv[3] = {coef[p] * (r[0] - v[p][0]), coef[p] * (r[1] - v[p][1]), coef[p] * (r[2] - v[p][2])};
      double fnv[3];
      cross3(nr, fv, fnv);
      const int gi = g.side.g[3 * f + p];
      const double2 xj = g.x[gi], xm = g.x[gi + g.N];
      for (int c = 0; c < 3; c++) {
        Ht[c][0] -= xj.x * fnv[c];
        Ht[c][1] -= xj.y * fnv[c];
        Et[c][0] += xm.x * fnv[c];
        Et[c][1] += xm.y * fnv[c];
      }
      double2 qe = make_double2(0.0, 0.0), qh = make_double2(0.0, 0.0);
      for (int c = 0; c < 3; c++) {
        qe.x += fv[c] * Ei[c][0];
        qe.y += fv[c] * Ei[c][1];
        qh.x += fv[c] * Hi[c][0];
        qh.y += fv[c] * Hi[c][1];
      }
      const double2 t1 = cm(g.lam[gi], qe), t2 = cm(g.lam[gi + g.N], qh);
      src.x += g.rule.w[t] * (t1.x + t2.x) * A;
      src.y += g.rule.w[t] * (t1.y + t2.y) * A;
    }
    double s1 = 0;
    for (int c = 0; c < 3; c++) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      const double re = (Et[c1][0] * Ht[c2][0] + Et[c1][1] * Ht[c2][1]) - (Et[c2][0] * Ht[c1][0] + Et[c2][1] * Ht[c1][1]);
      s1 += re * nr[c];
    }
    int1 += g.rule.w[t] * 0.5 * s1;
  }
  return g.cabs_c * int1 * A + src.x;
}

__global__ __launch_bounds__(128) void vertex_face_kernel(VFaceArgs g) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.F) return;
  VFace F0;
  VPairArgs dummy;  // vload_face needs only the node and side tables
  dummy.p = g.p;
  dummy.fnodes = g.fnodes;
  dummy.side = g.side;
  vload_face(dummy, f, F0);
  double L = 0;
  for (int v = 0; v < 3; v++) {
    double e[3];
    for (int k = 0; k < 3; k++) e[k] = F0.P[(v + 1) % 3][k] - F0.P[v][k];
    L += norm3(e);
  }
  const double h = 1e-3 * L / 3.0;
  const double steps[4] = {h, -h, 2.0 * h, -2.0 * h};
  for (int v = 0; v < 3; v++)
    for (int c = 0; c < 3; c++) {
      double val[4];
      for (int st = 0; st < 4; st++) {
        VFace Fs = F0;
        Fs.P[v][c] += steps[st];
        val[st] = vface_value(g, f, Fs);
      }
      g.fg[(int64_t)f * 9 + 3 * v + c] = (8.0 * (val[0] - val[1]) - (val[2] - val[3])) / (12.0 * h);
    }
}

int vertex_adjoint_fill(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t N,
                        const double2 *x, const double2 *lam, double *fg, int fd_only) {
  const cd I(0.0, 1.0);
  const cd k = ri * omega / kC0;
  const cd c1 = -(I * omega) * mu0();
  const cd c2 = -1.0 / ((I * omega) * kEps0 * (ri * ri));
  const cd zsi = (ri * ri) / (kEta0 * kEta0);
  const cd kk8 = (k * k) / (8.0 * MOM_PI);
  VPairArgs g;
  g.F = (int)m->F;
  g.rule = rule;
  g.p = m->p.as<double>();
  g.fnodes = m->fnodes.as<int>();
  g.side = sides(m, true);
  g.cp = m->cls();
  const double thr = m->avelen * 3;
  g.thr2 = thr * thr;
  g.kr = k.real();
  g.ki = k.imag();
  if (!(std::fabs(g.kr) * m->extent < kCisMaxArg))
    return fail(MOM_ERR_ARG, "Re(k) times the mesh extent is beyond the fill's phase range (2^19 rad)");
  g.c1 = make_double2(c1.real(), c1.imag());
  g.c2 = make_double2(c2.real(), c2.imag());
  g.zsi = make_double2(zsi.real(), zsi.imag());
  g.kk4 = make_double2(1.0 / (4.0 * MOM_PI), 0.0);
  g.kk8 = make_double2(kk8.real(), kk8.imag());
  g.x = x;
  g.lam = lam;
  g.N = N;
  g.fd_only = fd_only;
  g.pairs = m->near.as<int>();
  g.near_off = m->near_off.as<int64_t>();
  g.npairs = m->nnear;
  const int64_t F = m->F;
  g.ntx = (int)((F + kVTN - 1) / kVTN);
  g.nty = (int)((F + kVTM - 1) / kVTM);
  // tiles wholly below the diagonal write nothing: their slots stay zero
  DevBuf pt, ps, no;
  int rc;
  if ((rc = pt.alloc(sizeof(double) * 9 * F * g.ntx)) || (rc = ps.alloc(sizeof(double) * 9 * F * g.nty)) ||
      (!fd_only && m->nnear > 0 && (rc = no.alloc(sizeof(double) * 18 * m->nnear))))
    return rc;
  hipStream_t s = ctx->stream;
  MOM_HIP(hipMemsetAsync(pt.ptr, 0, pt.bytes, s));
  MOM_HIP(hipMemsetAsync(ps.ptr, 0, ps.bytes, s));
  if (no.ptr) MOM_HIP(hipMemsetAsync(no.ptr, 0, no.bytes, s));
  g.part_t = pt.as<double>();
  g.part_s = ps.as<double>();
  g.near_out = no.ptr ? no.as<double>() : nullptr;
  const dim3 grid((unsigned)g.ntx, (unsigned)g.nty);
  if (fd_only)
    hipLaunchKernelGGL(vertex_pair_kernel<true>, grid, dim3(128), 0, s, g);
  else
    hipLaunchKernelGGL(vertex_pair_kernel<false>, grid, dim3(128), 0, s, g);
  MOM_HIP(hipGetLastError());
  if (g.near_out) {
    hipLaunchKernelGGL(vertex_near_kernel, dim3((unsigned)((m->nnear + 127) / 128)), dim3(128), 0, s, g);
    MOM_HIP(hipGetLastError());
    const dim3 ga((unsigned)((9 * m->nnear + 127) / 128));
    hipLaunchKernelGGL(vertex_near_analytic_kernel<0>, ga, dim3(128), 0, s, g);
    hipLaunchKernelGGL(vertex_near_analytic_kernel<1>, ga, dim3(128), 0, s, g);
    hipLaunchKernelGGL(vertex_near_analytic_kernel<2>, ga, dim3(128), 0, s, g);
    MOM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(vertex_face_sum_kernel, dim3((unsigned)((F + 127) / 128)), dim3(128), 0, s, g, fg);
  MOM_HIP(hipGetLastError());
  // the partial sums must outlive the queued kernels
  MOM_HIP(hipStreamSynchronize(s));
  return 0;
}

int vertex_adjoint_faces(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t N,
                         const double2 *x, const double2 *lam, double th, double ph, double ps, double *fg) {
  VFaceArgs g;
  g.F = (int)m->F;
  g.rule = rule;
  g.p = m->p.as<double>();
  g.fnodes = m->fnodes.as<int>();
  g.side = sides(m, true);
  g.S.dir[0] = sin(th) * cos(ph);
  g.S.dir[1] = sin(th) * sin(ph);
  g.S.dir[2] = cos(th);
  g.S.pol[0] = cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph);
  g.S.pol[1] = -cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph);
  g.S.pol[2] = sin(ps) * sin(th);
  const cd k = ri * omega / kC0;
  const cd hfac = k / (omega * mu0());
  g.kr = k.real();
  g.ki = k.imag();
  g.hfac = make_double2(hfac.real(), hfac.imag());
  g.cabs_c = -2.0 / (kEps0 * kC0 * ri.real());
  g.x = x;
  g.lam = lam;
  g.N = N;
  g.fg = fg;
  hipLaunchKernelGGL(vertex_face_kernel, dim3((unsigned)((m->F + 127) / 128)), dim3(128), 0, ctx->stream, g);
  MOM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mom

extern "C" int mom_test_fill_lossless(int32_t enable) {
  if (enable != 0 && enable != 1) return mom::fail(MOM_ERR_ARG, "lossless switch must be 0 or 1");
  mom::g_fill_lossless = enable;
  return 0;
}
