// Internal declarations of libmomgpu (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mom_gpu.h"
#include "../../include/mom_gpu_testing.h"

namespace mom {

// constants.f90:8-16.  pi is a default-REAL literal in the reference, so its
// value is the float32 rounding of 3.141592653589 widened to double.
#define MOM_PI ((double)3.141592653589f)
constexpr double kEps0 = 8.8541878176e-12;
constexpr double kC0 = 2.997924580003e8;
constexpr double kEta0 = 3.767303134622e2;
inline double mu0() { return 4 * MOM_PI * 1e-7; }

constexpr int kMaxQ = 16;  // largest triangle rule supported by the kernels
constexpr int kSlab = 1024; // source faces per fill slab (face-side planes of 18*kSlab*F complex)

using cd = std::complex<double>;

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define MOM_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::mom::fail(MOM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define MOM_TRY(call)      \
  do {                     \
    int rc_ = (call);      \
    if (rc_) return rc_;   \
  } while (0)

// Device buffer with RAII.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  int alloc(size_t n);
  void release();
  template <class T> T *as() const { return static_cast<T *>(ptr); }
  void swap(DevBuf &o) {
    std::swap(ptr, o.ptr);
    std::swap(bytes, o.bytes);
  }
};

// Host <-> device copies on the context's stream through two pinned staging
// buffers: caller (pageable) memory is only read and written by the CPU, so
// the runtime never has to pin it.  copy_h2d returns once the source has
// been consumed (the device copy may still be in flight on the stream);
// copy_d2h returns when the destination holds the data.  The 2-D forms copy
// `rows` runs of `width` bytes with the given pitches.
int copy_h2d(mom_ctx *ctx, void *dev, const void *host, size_t bytes);
int copy_d2h(mom_ctx *ctx, void *host, const void *dev, size_t bytes);
int copy_h2d_2d(mom_ctx *ctx, void *dev, size_t dpitch, const void *host, size_t hpitch, size_t width,
                size_t rows);
int copy_d2h_2d(mom_ctx *ctx, void *host, size_t hpitch, const void *dev, size_t dpitch, size_t width,
                size_t rows);

}  // namespace mom

struct mom_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  mom::DevBuf qp;       // quadrature points of the last mesh/rule (scratch)
  mom::DevBuf scratch;  // general scratch
  mom::DevBuf panel_t;  // transposed U12 panel of the trailing updates (lu.hip)
  void *stage[2] = {nullptr, nullptr};   // pinned staging buffers
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // second stream of the LU look-ahead (panel k+1 factored while the
  // trailing update of step k runs on `stream`), created on first use, with
  // its own GEMM scratch; fork/join events between the two
  hipStream_t aux = nullptr;
  mom::DevBuf panel_t2;
  mom::DevBuf ublk;  // E = I - inv(L11) of the last factored panel (complex64, NB x NB; lu.hip)
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // fill autotuner: source faces per wave of the pair kernel chosen by timing
  // the candidates on the first slab of the first fill (0: not yet tuned)
  int fill_spw = 0;
  double fill_spw_ms[3] = {0, 0, 0};  // measured slab times of the candidates 2, 4, 8
  // event timing of the fill kernels (mom_ctx_fill_profile): when on, every
  // pair-kernel and assembly launch is bracketed by events and the totals
  // below accumulate: [0] pair ms, [1] pair launches, [2] pair point pairs
  // (test x source quadrature points), [3] assembly ms, [4] assembly
  // launches, [5] near-part ms
  int fill_profile = 0;
  double fill_prof[6] = {0, 0, 0, 0, 0, 0};
};

// Device copy of one domain mesh, with per-face-side RWG data precomputed.
struct mom_mesh {
  mom_ctx *ctx = nullptr;
  int64_t V = 0, F = 0, E = 0;
  double avelen = 0;
  double extent = 0;  // diagonal of the nodes' bounding box (bounds every point distance)
  double rmax = 0;    // largest distance of a node from the origin (bounds |J r - r'| <= 2 rmax for images)
  int64_t nnear = 0;
  int64_t max_parent = -1;  // largest global index among the mesh's edges
  std::vector<double> h_fcp;  // host centroids (near list)
  mom::DevBuf p, fnodes, fn, fm, fs, farea, fcp;
  // near/far classification by another mesh's centroids (and avelen, which
  // only classifies), set by mom_mesh_classify_like; empty: its own
  std::vector<double> h_ccls;
  mom::DevBuf ccls;
  const std::vector<double> &h_cls() const { return h_ccls.empty() ? h_fcp : h_ccls; }
  const double *cls() const { return ccls.ptr ? ccls.as<double>() : fcp.as<double>(); }
  // per face side s = 3*f + e:
  mom::DevBuf s_g;     // int32 parent (global) index of the edge
  mom::DevBuf s_l;     // int32 local edge index
  mom::DevBuf s_coef;  // sign * L/(2A): f(r) = coef * (r - v)
  mom::DevBuf s_div;   // sign * L / A (surface divergence)
  mom::DevBuf s_v;     // free vertex (3 doubles)
  mom::DevBuf near;    // int32 pairs (test face, source face), sorted
  mom::DevBuf near_off;  // int64 F + 1: the pairs of test face m are near_off[m] .. near_off[m+1]
  // fill assembly: face-side codes 3f+s of each edge's two faces (-1 none),
  // and per slab of kSlab source faces the edges with a side in the slab
  mom::DevBuf eside, slab_edges;
  std::vector<int> h_slab_ptr;
  std::vector<int> h_eparent;
  // host connectivity (face edges, edge faces, edge face-side codes)
  std::vector<int32_t> h_fedges, h_efaces, h_eside;
  std::vector<int32_t> h_fnodes;  // [F][3] (node gradients are summed through it on the host)
};

namespace mom {

struct QuadRule {
  int Q;
  double w[kMaxQ];
  double nodes[3 * kMaxQ];  // (Q,3) column-major, as quad_data%nodes
};
int make_rule(int32_t nq, const double *qw, const double *qnodes, QuadRule *r);

// Representation matrices of a mirror-symmetry group (sysmat.f90:56-101
// with bc.f90:170-193): the image of a source face under the group element
// adds e * (its block) to every representation's matrix, whose rows and
// columns are the unknowns that the representation's boundary conditions
// keep, numbered in order (map: unknown u in [0, 2 Noff) -> row, -1 removed).
constexpr int kMaxReps = 8;
struct RepOut {
  double2 *Z[kMaxReps];
  int64_t ld[kMaxReps];
  const int *map[kMaxReps];
  double2 e[kMaxReps];  // field action of the element in each representation (complex for rotations)
};

// Output of the fill kernels.
struct FillTarget {
  double2 *Z = nullptr;  // column-major complex matrix
  int64_t ld = 0;
  int64_t Noff = 0;   // offset of the M rows/columns (nedges_total)
  int mode = 0;       // 0: PMCHWT blocks with global indices; 1/2/3: D/H/K, local indices
  int local = 0;      // mode 0 with local indices (the domain's own 2E x 2E block, Noff = E)
  int dk = 0;         // the k-derivative of the operators (same prefactors c1, c2, zsi)
  double2 zsi;        // Zsi = ri^2 / eta0^2 (mode 0)
  int nrep = 0;       // > 0: mode 0 into the representation matrices in `rep` instead of Z
  double detj = 1.0;  // determinant of the element's Jacobian (nrep > 0)
  RepOut rep;
};

// A symmetry group (symmetry.f90:26-241): products of the mirrors of the
// planes x = 0, y = 0, z = 0 and of rotations about the z axis.  Per
// element its Jacobian J (row-major, J[3 i + j] = J_ij), determinant,
// identifier and generator bits, and the field action
// ef[element][representation] (the conjugated representation; complex for
// rotations).
struct Group {
  int n = 1;
  double J[kMaxReps][9];
  double detj[kMaxReps];
  double2 ef[kMaxReps][kMaxReps];
  int id[kMaxReps];
  int genbits[kMaxReps];
};
constexpr int kGidIdentity = 1, kGidMxp = 2, kGidMyp = 3, kGidMzp = 4, kGidRz = 5;  // symmetry.f90:20-24
constexpr int kBndNone = 1, kBndXplane = 2, kBndYplane = 3, kBndZplane = 4;  // mesh.f90:16-25
// generator codes of the `symm` line: 1, 2, 3 = mxp, myp, mzp; kGenRz + n = rn
constexpr int kGenRz = 100;
Group identity_group();
// read_symm's product of the generators gens[i] (codes above)
int symm_group(int ngen, const int32_t *gens, Group *g);
// whether J is diagonal (a product of mirrors) / its own inverse
bool jac_diagonal(const double *J);
bool jac_involution(const double *J);

// An image J of the source faces (symmetry.f90:61-156): the fill integrates
// the test faces moved by M = J^T against the source faces, which equals
// the reference's integral over the source moved by J (srcint.f90:155-224,
// 368-448), J being orthogonal.  For a mirror (diagonal J, its own inverse)
// each unordered face pair is integrated once as without an image; a
// rotation's transposed blocks belong to its inverse element, so its
// ordered pairs are integrated one by one.
struct Image {
  double M[9];  // J^T, row-major
  double detj;
  bool diag;    // J (and M) diagonal: the mirror code path
  bool invol;   // J = J^-1: the transposed blocks are this element's own
};

// api.hip: near face pairs (test, source) sorted, of the mesh with itself or,
// with J (row-major), of the test faces against the source faces moved by J:
// |c_m - J c_n|^2 < (3 avelen)^2 in the reference's order of operations
void near_pairs(const std::vector<double> &cp, int64_t F, double avelen, std::vector<int32_t> &out,
                const double *J = nullptr);
// c - J c' for the near rule of an image (no contraction), shared by the host
// list and the device kernels so that both classify every pair alike
__host__ __device__ inline void image_near_diff(const double *J, const double *c, const double *cn, double *d) {
#pragma clang fp contract(off)
  for (int i = 0; i < 3; i++) {
    double t = J[3 * i] * cn[0];
    t = t + J[3 * i + 1] * cn[1];
    t = t + J[3 * i + 2] * cn[2];
    d[i] = c[i] - t;
  }
}
// off[m] .. off[m+1]: the pairs of test face m in a list sorted by test face (F + 1 entries)
void near_offsets(const std::vector<int32_t> &pairs, int64_t F, std::vector<int64_t> &off);

// fill.hip
int compute_quad_points(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule);
int fill_domain(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                const FillTarget &t, const Image *img = nullptr);
// q[:, s * G.n + r]: the plane wave s projected on representation r (identity group: one column per source)
int srcvec_pw(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
              int64_t nsrc, const double *angles, int64_t Nt, double2 *q, int64_t ldq,
              const Group &G = identity_group());
int fill_system(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *ri,
                int64_t nedges_total, double omega, int32_t nq, const double *qw,
                const double *qnodes, double2 *Z, int64_t ldz);
// x: one column of 2 Nt per representation of G (leading dimension ldx)
int cs_partials(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                const double2 *x, int64_t Nt, double th, double ph, double ps, double *partial,
                const Group &G = identity_group(), int64_t ldx = 0);
// (Nt, Np, Lt, Lp) of far_fields (ffields.f90:16-122) for ndir directions,
// x the device solution [2 Nt], k the medium's wavenumber
size_t far_work_bytes(const mom_mesh *m, int Q, const Group &G, int64_t ndir);
int far_fields_device(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, const double2 *x,
                      int64_t Nt, double r, int64_t ndir, const double *theta, const double *phi, double *out,
                      void *work, const Group &G = identity_group(), int64_t ldx = 0);
// the allocation-free far field: scratch in a fixed device array set up per
// device by no_alloc_init (called when a context is created)
int no_alloc_init(int device);
int far_fields_no_alloc(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri,
                        const double2 *x, int64_t Nt, double r, int64_t ndir, const double *theta, const double *phi,
                        double *out);
// sysmat_pmchwt for a mirror group (symm.hip): every representation r's
// reduced matrix Z[r] (order[r] = kept unknowns, padded to even by a unit
// diagonal), bnd[N] the boundary class of each master edge
int fill_system_symm(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *ri, int64_t N,
                     const int32_t *bnd, const Group &G, double omega, int32_t nq, const double *qw,
                     const double *qnodes, double2 *const *Z, const int64_t *ld, const int64_t *order);
// Vertex-coordinate adjoint (fill.hip), fg[F][9] (device) written per face,
// Re d/d(coordinate c of the face's local node v) at 3 v + c:
// Re d/dP of lambda^T Z_d x for the domain's blocks (far pairs in closed
// form, near pairs' smoothed part in closed form and analytic part by
// central differences, frozen near/far classification) ...
// (fd_only: every pair by differences, as a check of the closed forms)
int vertex_adjoint_fill(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t N,
                        const double2 *x, const double2 *lam, double *fg, int fd_only = 0);
// ... and of sigma_abs|_x + Re(lambda^T q) for the exterior domain and one plane wave
int vertex_adjoint_faces(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t N,
                         const double2 *x, const double2 *lam, double th, double ph, double ps, double *fg);
// Transpose of the far-field map (fill.hip): g[2 Nt] (host) = L^T v for
// adjoint values v[8 ndir] of (et, ep, ht, hp) per direction (host arrays)
int far_fields_adjoint(mom_ctx *ctx, const mom_mesh *m, const QuadRule &rule, double omega, cd ri, int64_t Nt,
                       double r, int64_t ndir, const double *theta, const double *phi, const double *v, double *g);
// Per-side pairs w[3F][2] (device) summed per edge into the global device
// vector g[2 Nt] at the edges' parents (rows parent, parent + Nt), face order
int gather_sides(mom_ctx *ctx, const mom_mesh *m, const double2 *w, int64_t Nt, double2 *g);
// y = M x (device, complex128, M nr x nc with leading dimension ld), in
// column chunks spread over the device (adjoint.hip)
int matvec(mom_ctx *ctx, const double2 *M, int64_t ld, int64_t nr, int64_t nc, const double2 *x, double2 *y);
// sysmat_pmchwt_nls (sysmat.f90:112-229) into the device matrix Z, with the
// products of the nonlinear domains' blocks and src_coef subtracted from the
// host sources src (nls.hip)
int fill_system_nls(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *shri,
                    int64_t nedges_total, double omega, int32_t nq, const double *qw, const double *qnodes,
                    double2 *Z, int64_t ldz, int64_t nsrc, const double *src_coef, int64_t ldsc, double *src,
                    int64_t ldsrc);

}  // namespace mom
