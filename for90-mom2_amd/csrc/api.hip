// C-ABI entry points of libmomgpu: contexts, meshes, fill, RHS, post.
// Factorisation and solve entry points live in lu.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "mom_internal.hpp"

namespace mom {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return fail(MOM_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed: " +
                                   hipGetErrorString(e));
  }
  bytes = n;
  return 0;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

constexpr size_t kStageBytes = size_t(16) << 20;

static int stage_init(mom_ctx *ctx) {
  for (int k = 0; k < 2; k++) {
    if (!ctx->stage[k]) MOM_HIP(hipHostMalloc(&ctx->stage[k], kStageBytes, hipHostMallocDefault));
    if (!ctx->stage_ev[k]) MOM_HIP(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
  }
  return 0;
}

int copy_h2d_2d(mom_ctx *ctx, void *dev, size_t dpitch, const void *host, size_t hpitch, size_t width,
                size_t rows) {
  if (!width || !rows) return 0;
  int rc = stage_init(ctx);
  if (rc) return rc;
  const char *h = static_cast<const char *>(host);
  char *d = static_cast<char *>(dev);
  if (width > kStageBytes) {  // long runs: one run at a time in stage-sized pieces
    for (size_t r = 0; r < rows; r++)
      if ((rc = copy_h2d(ctx, d + r * dpitch, h + r * hpitch, width))) return rc;
    return 0;
  }
  const size_t per = kStageBytes / width;
  int k = 0;
  for (size_t r0 = 0; r0 < rows; r0 += per, k ^= 1) {
    const size_t nr = std::min(per, rows - r0);
    MOM_HIP(hipEventSynchronize(ctx->stage_ev[k]));
    char *st = static_cast<char *>(ctx->stage[k]);
    for (size_t r = 0; r < nr; r++) std::memcpy(st + r * width, h + (r0 + r) * hpitch, width);
    MOM_HIP(hipMemcpy2DAsync(d + r0 * dpitch, dpitch, st, width, width, nr, hipMemcpyHostToDevice, ctx->stream));
    MOM_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
  }
  return 0;
}

int copy_h2d(mom_ctx *ctx, void *dev, const void *host, size_t bytes) {
  if (!bytes) return 0;
  int rc = stage_init(ctx);
  if (rc) return rc;
  int k = 0;
  for (size_t off = 0; off < bytes; off += kStageBytes, k ^= 1) {
    const size_t len = std::min(kStageBytes, bytes - off);
    MOM_HIP(hipEventSynchronize(ctx->stage_ev[k]));
    std::memcpy(ctx->stage[k], static_cast<const char *>(host) + off, len);
    MOM_HIP(hipMemcpyAsync(static_cast<char *>(dev) + off, ctx->stage[k], len, hipMemcpyHostToDevice, ctx->stream));
    MOM_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
  }
  return 0;
}

int copy_d2h_2d(mom_ctx *ctx, void *host, size_t hpitch, const void *dev, size_t dpitch, size_t width,
                size_t rows) {
  if (!width || !rows) return 0;
  int rc = stage_init(ctx);
  if (rc) return rc;
  char *h = static_cast<char *>(host);
  const char *d = static_cast<const char *>(dev);
  if (width > kStageBytes) {
    for (size_t r = 0; r < rows; r++)
      if ((rc = copy_d2h(ctx, h + r * hpitch, d + r * dpitch, width))) return rc;
    return 0;
  }
  const size_t per = kStageBytes / width;
  const size_t nchunk = (rows + per - 1) / per;
  auto issue = [&](size_t c) -> int {
    const int k = (int)(c & 1);
    const size_t r0 = c * per, nr = std::min(per, rows - r0);
    MOM_HIP(hipEventSynchronize(ctx->stage_ev[k]));
    MOM_HIP(hipMemcpy2DAsync(ctx->stage[k], width, d + r0 * dpitch, dpitch, width, nr, hipMemcpyDeviceToHost,
                             ctx->stream));
    MOM_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
    return 0;
  };
  if ((rc = issue(0))) return rc;
  if (nchunk > 1 && (rc = issue(1))) return rc;
  for (size_t c = 0; c < nchunk; c++) {
    const int k = (int)(c & 1);
    const size_t r0 = c * per, nr = std::min(per, rows - r0);
    MOM_HIP(hipEventSynchronize(ctx->stage_ev[k]));
    const char *st = static_cast<const char *>(ctx->stage[k]);
    for (size_t r = 0; r < nr; r++) std::memcpy(h + (r0 + r) * hpitch, st + r * width, width);
    if (c + 2 < nchunk && (rc = issue(c + 2))) return rc;
  }
  return 0;
}

int copy_d2h(mom_ctx *ctx, void *host, const void *dev, size_t bytes) {
  if (!bytes) return 0;
  const size_t rows = (bytes + kStageBytes - 1) / kStageBytes;
  if (rows == 1) return copy_d2h_2d(ctx, host, bytes, dev, bytes, bytes, 1);
  // whole stage-sized pieces, then the tail
  const size_t whole = bytes / kStageBytes;
  int rc = copy_d2h_2d(ctx, host, kStageBytes, dev, kStageBytes, kStageBytes, whole);
  if (rc || bytes == whole * kStageBytes) return rc;
  const size_t off = whole * kStageBytes;
  return copy_d2h_2d(ctx, static_cast<char *>(host) + off, bytes - off, static_cast<const char *>(dev) + off,
                     bytes - off, bytes - off, 1);
}

template <class T>
static int upload(mom_ctx *ctx, DevBuf &b, const T *src, size_t count) {
  int rc = b.alloc(count * sizeof(T));
  if (rc) return rc;
  return copy_h2d(ctx, b.ptr, src, count * sizeof(T));
}

int make_rule(int32_t nq, const double *qw, const double *qnodes, QuadRule *r) {
  if (nq < 1 || nq > kMaxQ || !qw || !qnodes)
    return fail(MOM_ERR_ARG, "triangle rule must have 1.." + std::to_string(kMaxQ) + " points");
  r->Q = nq;
  for (int i = 0; i < kMaxQ; i++) r->w[i] = i < nq ? qw[i] : 0.0;
  for (int i = 0; i < 3 * kMaxQ; i++) r->nodes[i] = i < 3 * nq ? qnodes[i] : 0.0;
  return 0;
}

// Near face pairs: centroid distance^2 < (3*avelen)^2 (sysmat.f90:233-248),
// evaluated exactly as the reference does, found with a uniform grid.  With
// J, the pairs of the test faces and the source faces moved by J:
// |c_m - J c_n|^2 in the reference's order of operations
// (image_near_diff), the candidates taken around J^T c_m.
void near_pairs(const std::vector<double> &cp, int64_t F, double avelen, std::vector<int32_t> &out,
                const double *J) {
#pragma clang fp contract(off)
  const double thr = avelen * 3;
  const double thr2 = thr * thr;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t f = 0; f < F; f++)
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min(lo[c], cp[3 * f + c]);
      hi[c] = std::max(hi[c], cp[3 * f + c]);
    }
  const double cell = thr > 0 ? thr : 1.0;
  int64_t dim[3];
  for (int c = 0; c < 3; c++) dim[c] = std::max<int64_t>(1, (int64_t)((hi[c] - lo[c]) / cell) + 1);
  auto key = [&](int64_t i, int64_t j, int64_t k) { return (i * dim[1] + j) * dim[2] + k; };
  std::unordered_map<int64_t, std::vector<int32_t>> grid;
  std::vector<int64_t> cell_of(3 * F);
  for (int64_t f = 0; f < F; f++) {
    int64_t ic[3];
    for (int c = 0; c < 3; c++) ic[c] = std::min(dim[c] - 1, (int64_t)((cp[3 * f + c] - lo[c]) / cell));
    for (int c = 0; c < 3; c++) cell_of[3 * f + c] = ic[c];
    grid[key(ic[0], ic[1], ic[2])].push_back((int32_t)f);
  }
  std::vector<int32_t> nb;
  for (int64_t m = 0; m < F; m++) {
    nb.clear();
    double tc[3];
    int64_t tcell[3];
    for (int c = 0; c < 3; c++) {
      tc[c] = J ? J[c] * cp[3 * m] + J[3 + c] * cp[3 * m + 1] + J[6 + c] * cp[3 * m + 2] : cp[3 * m + c];
      // a moved centroid may lie outside the grid: cells beyond it are skipped below
      const double u = (tc[c] - lo[c]) / cell;
      tcell[c] = J ? (u < -2.0 ? -3 : std::min(dim[c] + 1, (int64_t)std::floor(u))) : cell_of[3 * m + c];
    }
    for (int64_t a = -1; a <= 1; a++)
      for (int64_t b = -1; b <= 1; b++)
        for (int64_t c = -1; c <= 1; c++) {
          int64_t i = tcell[0] + a, j = tcell[1] + b, k = tcell[2] + c;
          if (i < 0 || j < 0 || k < 0 || i >= dim[0] || j >= dim[1] || k >= dim[2]) continue;
          auto it = grid.find(key(i, j, k));
          if (it == grid.end()) continue;
          for (int32_t n : it->second) {
            double d[3];
            if (J) {
              image_near_diff(J, &cp[3 * m], &cp[3 * n], d);
            } else {
              d[0] = tc[0] - cp[3 * n];
              d[1] = tc[1] - cp[3 * n + 1];
              d[2] = tc[2] - cp[3 * n + 2];
            }
            double ds = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (ds < thr2) nb.push_back(n);
          }
        }
    std::sort(nb.begin(), nb.end());
    for (int32_t n : nb) {
      out.push_back((int32_t)m);
      out.push_back(n);
    }
  }
}

void near_offsets(const std::vector<int32_t> &pairs, int64_t F, std::vector<int64_t> &off) {
  off.assign(F + 1, 0);
  const int64_t np = (int64_t)pairs.size() / 2;
  for (int64_t k = 0; k < np; k++) off[pairs[2 * k] + 1]++;
  for (int64_t f = 0; f < F; f++) off[f + 1] += off[f];
}

}  // namespace mom

using namespace mom;

extern "C" {

const char *mom_last_error(void) { return g_err.c_str(); }

int mom_ctx_create(int32_t device, mom_ctx **ctx) {
  if (!ctx) return fail(MOM_ERR_ARG, "ctx is NULL");
  int ndev = 0;
  MOM_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MOM_ERR_ARG, "no such device");
  MOM_HIP(hipSetDevice(device));
  int rc = no_alloc_init(device);
  if (rc) return rc;
  mom_ctx *c = new mom_ctx;
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(MOM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *ctx = c;
  return 0;
}

int mom_ctx_fill_tuning(const mom_ctx *ctx, int32_t *spw, double *ms) {
  if (!ctx || !spw) return fail(MOM_ERR_ARG, "NULL argument");
  *spw = ctx->fill_spw;
  if (ms)
    for (int c = 0; c < 3; c++) ms[c] = ctx->fill_spw_ms[c];
  return 0;
}

int mom_test_fill_spw(mom_ctx *ctx, int32_t spw) {
  if (!ctx) return fail(MOM_ERR_ARG, "NULL argument");
  if (spw != 0 && spw != 2 && spw != 4 && spw != 8) return fail(MOM_ERR_ARG, "sources per wave must be 0, 2, 4 or 8");
  ctx->fill_spw = spw;
  return 0;
}

int mom_ctx_fill_profile(mom_ctx *ctx, int32_t enable, double *out) {
  if (!ctx) return fail(MOM_ERR_ARG, "NULL argument");
  if (out)
    for (int c = 0; c < 6; c++) out[c] = ctx->fill_prof[c];
  if (enable >= 0) {
    ctx->fill_profile = enable ? 1 : 0;
    for (int c = 0; c < 6; c++) ctx->fill_prof[c] = 0;
  }
  return 0;
}

int mom_ctx_destroy(mom_ctx *ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->qp.release();
  ctx->scratch.release();
  if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
  ctx->panel_t.release();
  ctx->panel_t2.release();
  ctx->ublk.release();
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
  for (int k = 0; k < 2; k++) {
    if (ctx->stage[k]) (void)hipHostFree(ctx->stage[k]);
    if (ctx->stage_ev[k]) (void)hipEventDestroy(ctx->stage_ev[k]);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int mom_mesh_create(mom_ctx *ctx, int64_t nnodes, const double *p, int64_t nfaces,
                    const int32_t *fnodes, const int32_t *fedges, const double *fn,
                    const double *fcp, const double *fs, const double *fm, const double *farea,
                    int64_t nedges, const int32_t *efaces, const int32_t *ebn, const double *elen,
                    const int32_t *eparent, double avelen, mom_mesh **mesh) {
  if (!ctx || !mesh || !p || !fnodes || !fedges || !fn || !fcp || !fs || !fm || !farea ||
      !efaces || !ebn || !elen || !eparent)
    return fail(MOM_ERR_ARG, "NULL argument to mom_mesh_create");
  if (nnodes < 3 || nfaces < 1 || nedges < 1 || nfaces > (1 << 28) || nedges > (1 << 29))
    return fail(MOM_ERR_ARG, "mesh sizes out of range");
  if (!(avelen > 0) || !std::isfinite(avelen)) return fail(MOM_ERR_ARG, "avelen must be > 0");
  const int64_t F = nfaces, E = nedges;
  for (int64_t i = 0; i < 3 * F; i++) {
    if (fnodes[i] < 0 || fnodes[i] >= nnodes) return fail(MOM_ERR_ARG, "face node index out of range");
    if (fedges[i] < 0 || fedges[i] >= E) return fail(MOM_ERR_ARG, "face edge index out of range");
  }
  // Per face side: sign (get_face_sign, mesh.f90:1326-1338), coefficient
  // L/(2A), divergence sign*L/A (rwgf.f90:34-78) and free vertex
  // (get_face_bnode, mesh.f90:1358-1369).
  std::vector<int32_t> sg(3 * F), sl(3 * F);
  std::vector<double> scoef(3 * F), sdiv(3 * F), sv(9 * F);
  for (int64_t f = 0; f < F; f++)
    for (int e = 0; e < 3; e++) {
      const int32_t ge = fedges[3 * f + e];
      double sign;
      int32_t bn;
      if (efaces[2 * ge] == f) {
        sign = 1.0;
        bn = ebn[2 * ge];
      } else if (efaces[2 * ge + 1] == f) {
        sign = -1.0;
        bn = ebn[2 * ge + 1];
      } else {
        return fail(MOM_ERR_ARG, "corrupted mesh connectivity: face " + std::to_string(f) +
                                     " is neither face of its edge " + std::to_string(ge));
      }
      if (bn < 0 || bn >= nnodes) return fail(MOM_ERR_ARG, "free vertex index out of range");
      if (eparent[ge] < 0) return fail(MOM_ERR_ARG, "edge without parent index");
      const double L = elen[ge], A = farea[f];
      sg[3 * f + e] = eparent[ge];
      sl[3 * f + e] = ge;
      scoef[3 * f + e] = sign * (L / (2.0 * A));
      sdiv[3 * f + e] = sign * L / A;
      for (int c = 0; c < 3; c++) sv[3 * (3 * f + e) + c] = p[3 * bn + c];
    }
  mom_mesh *m = new mom_mesh;
  m->ctx = ctx;
  for (int64_t e = 0; e < E; e++) m->max_parent = std::max<int64_t>(m->max_parent, eparent[e]);
  m->V = nnodes;
  m->F = F;
  m->E = E;
  m->avelen = avelen;
  {
    double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    for (int64_t v = 0; v < nnodes; v++)
      for (int c = 0; c < 3; c++) {
        lo[c] = std::min(lo[c], p[3 * v + c]);
        hi[c] = std::max(hi[c], p[3 * v + c]);
      }
    m->extent = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                          (hi[2] - lo[2]) * (hi[2] - lo[2]));
    double r2 = 0;
    for (int64_t v = 0; v < nnodes; v++)
      r2 = std::max(r2, p[3 * v] * p[3 * v] + p[3 * v + 1] * p[3 * v + 1] + p[3 * v + 2] * p[3 * v + 2]);
    m->rmax = std::sqrt(r2);
  }
  m->h_fcp.assign(fcp, fcp + 3 * F);
  // face-side codes of each edge and the slab edge lists of the fill
  std::vector<int32_t> eside(2 * E, -1);
  for (int64_t f = 0; f < F; f++)
    for (int e = 0; e < 3; e++) {
      const int32_t ge = fedges[3 * f + e];
      const int slot = (efaces[2 * ge] == f) ? 0 : 1;
      eside[2 * ge + slot] = (int32_t)(3 * f + e);
    }
  std::vector<int32_t> slab_edges;
  m->h_slab_ptr.assign(1, 0);
  {
    std::vector<int64_t> seen(E, -1);
    const int64_t nslab = (F + kSlab - 1) / kSlab;
    for (int64_t k = 0; k < nslab; k++) {
      for (int64_t f = k * kSlab; f < std::min<int64_t>(F, (k + 1) * kSlab); f++)
        for (int e = 0; e < 3; e++) {
          const int32_t ge = fedges[3 * f + e];
          if (seen[ge] != k) {
            seen[ge] = k;
            slab_edges.push_back(ge);
          }
        }
      m->h_slab_ptr.push_back((int)slab_edges.size());
    }
  }
  m->h_eparent.assign(eparent, eparent + E);
  m->h_fnodes.assign(fnodes, fnodes + 3 * F);
  m->h_fedges.assign(fedges, fedges + 3 * F);
  m->h_efaces.assign(efaces, efaces + 2 * E);
  m->h_eside = eside;
  std::vector<int32_t> np;
  near_pairs(m->h_fcp, F, avelen, np);
  m->nnear = (int64_t)np.size() / 2;
  std::vector<int64_t> noff;
  near_offsets(np, F, noff);
  hipStream_t s = ctx->stream;
  int rc = 0;
  (void)hipSetDevice(ctx->device);
  if ((rc = upload(ctx, m->p, p, 3 * nnodes)) || (rc = upload(ctx, m->fnodes, fnodes, 3 * F)) ||
      (rc = upload(ctx, m->fn, fn, 3 * F)) || (rc = upload(ctx, m->fm, fm, 9 * F)) ||
      (rc = upload(ctx, m->fs, fs, 9 * F)) || (rc = upload(ctx, m->farea, farea, F)) ||
      (rc = upload(ctx, m->fcp, fcp, 3 * F)) || (rc = upload(ctx, m->s_g, sg.data(), 3 * F)) ||
      (rc = upload(ctx, m->s_l, sl.data(), 3 * F)) || (rc = upload(ctx, m->s_coef, scoef.data(), 3 * F)) ||
      (rc = upload(ctx, m->s_div, sdiv.data(), 3 * F)) || (rc = upload(ctx, m->s_v, sv.data(), 9 * F)) ||
      (rc = upload(ctx, m->near, np.data(), np.size())) || (rc = upload(ctx, m->near_off, noff.data(), noff.size())) ||
      (rc = upload(ctx, m->eside, eside.data(), eside.size())) ||
      (rc = upload(ctx, m->slab_edges, slab_edges.data(), slab_edges.size()))) {
    delete m;
    return rc;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete m;
    return fail(MOM_ERR_HIP, std::string("mesh upload: ") + hipGetErrorString(e));
  }
  *mesh = m;
  return 0;
}

int mom_mesh_destroy(mom_mesh *mesh) {
  if (!mesh) return 0;
  (void)hipSetDevice(mesh->ctx->device);
  (void)hipStreamSynchronize(mesh->ctx->stream);
  delete mesh;
  return 0;
}

int64_t mom_mesh_near_pairs(const mom_mesh *mesh) { return mesh ? mesh->nnear : -1; }

int mom_mesh_classify_like(mom_mesh *mesh, const mom_mesh *from) {
  if (!mesh || !from) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != from->ctx) return fail(MOM_ERR_ARG, "classify_like: the meshes belong to different contexts");
  if (mesh->F != from->F) return fail(MOM_ERR_ARG, "classify_like: the meshes have different face counts");
  mom_ctx *ctx = mesh->ctx;
  const int64_t F = mesh->F;
  (void)hipSetDevice(ctx->device);
  const std::vector<double> cls = from->h_cls();  // a copy: mesh may be from
  const double avelen = from->avelen;
  std::vector<int32_t> np;
  near_pairs(cls, F, avelen, np);
  std::vector<int64_t> noff;
  near_offsets(np, F, noff);
  DevBuf dcls, dnear, doff;
  int rc = 0;
  if ((rc = upload(ctx, dcls, cls.data(), cls.size())) || (rc = upload(ctx, dnear, np.data(), np.size())) ||
      (rc = upload(ctx, doff, noff.data(), noff.size())))
    return rc;
  // fills queued on the stream may still read the old lists
  MOM_HIP(hipStreamSynchronize(ctx->stream));
  mesh->ccls.swap(dcls);
  mesh->near.swap(dnear);
  mesh->near_off.swap(doff);
  mesh->h_ccls = cls;
  mesh->avelen = avelen;
  mesh->nnear = (int64_t)np.size() / 2;
  return 0;
}

int mom_compute_operator(mom_ctx *ctx, const mom_mesh *mesh, char op, double omega, double ri_re,
                         double ri_im, int32_t nq, const double *qw, const double *qnodes,
                         double *M, int64_t ldm) {
  if (!ctx || !mesh || !M) return fail(MOM_ERR_ARG, "NULL argument");
  int mode = op == 'D' ? 1 : op == 'H' ? 2 : op == 'K' ? 3 : 0;
  if (!mode) return fail(MOM_ERR_ARG, "op must be 'D', 'H' or 'K'");
  if (ldm < mesh->E) return fail(MOM_ERR_ARG, "ldm < nedges");
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const int64_t E = mesh->E;
  DevBuf dM;
  if ((rc = dM.alloc(sizeof(double2) * E * E))) return rc;
  MOM_HIP(hipMemsetAsync(dM.ptr, 0, sizeof(double2) * E * E, ctx->stream));
  FillTarget t;
  t.Z = dM.as<double2>();
  t.ld = E;
  t.mode = mode;
  if ((rc = fill_domain(ctx, mesh, rule, omega, cd(ri_re, ri_im), t))) return rc;
  return copy_d2h_2d(ctx, M, ldm * sizeof(double2), dM.ptr, E * sizeof(double2), E * sizeof(double2), E);
  return 0;
}

int mom_srcvec_pw(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega,
                  double ri_re, double ri_im, int64_t nsrc, const double *angles, int32_t nq,
                  const double *qw, const double *qnodes, double *q, int64_t ldq) {
  if (!ctx || !mesh || !angles || !q) return fail(MOM_ERR_ARG, "NULL argument");
  if (nsrc < 1 || nedges_total < 1 || ldq < 2 * nedges_total) return fail(MOM_ERR_ARG, "bad sizes");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const int64_t N2 = 2 * nedges_total;
  DevBuf dq;
  if ((rc = dq.alloc(sizeof(double2) * N2 * nsrc))) return rc;
  MOM_HIP(hipMemsetAsync(dq.ptr, 0, sizeof(double2) * N2 * nsrc, ctx->stream));
  if ((rc = srcvec_pw(ctx, mesh, rule, omega, cd(ri_re, ri_im), nsrc, angles, nedges_total,
                      dq.as<double2>(), N2)))
    return rc;
  return copy_d2h_2d(ctx, q, ldq * sizeof(double2), dq.ptr, N2 * sizeof(double2), N2 * sizeof(double2), nsrc);
  return 0;
}

static int cs_impl(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                   double ri_im, const double *x, double theta, double phi, double psi, int32_t nq, const double *qw,
                   const double *qnodes, double *csca, double *cabs, const Group &G) {
  if (!ctx || !mesh || !x || !csca || !cabs) return fail(MOM_ERR_ARG, "NULL argument");
  if (!(ri_re > 0)) return fail(MOM_ERR_ARG, "Re(ri) must be > 0");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "nedges_total too small");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const int64_t N2 = 2 * nedges_total;
  DevBuf dx, dpart;
  if ((rc = dx.alloc(sizeof(double2) * N2 * G.n)) || (rc = dpart.alloc(sizeof(double) * 2 * mesh->F))) return rc;
  if ((rc = copy_h2d(ctx, dx.ptr, x, sizeof(double2) * N2 * G.n))) return rc;
  if ((rc = cs_partials(ctx, mesh, rule, omega, cd(ri_re, ri_im), dx.as<double2>(), nedges_total, theta,
                        phi, psi, dpart.as<double>(), G, N2)))
    return rc;
  std::vector<double> part(2 * mesh->F);
  if ((rc = copy_d2h(ctx, part.data(), dpart.ptr, sizeof(double) * 2 * mesh->F))) return rc;
  // cs.f90:38, 97-102: sums in face order, then c = 2/(eps0 c0 Re ri)
  double sa = 0, ss = 0;
  for (int64_t f = 0; f < mesh->F; f++) {
    sa += part[2 * f];
    ss += part[2 * f + 1];
  }
  const double c = 2.0 / (kEps0 * kC0 * ri_re);
  *cabs = -c * sa;
  *csca = c * ss;
  return 0;
}

int mom_cs_prtsrf(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                  double ri_im, const double *x, double theta, double phi, double psi, int32_t nq, const double *qw,
                  const double *qnodes, double *csca, double *cabs) {
  return cs_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, theta, phi, psi, nq, qw, qnodes, csca, cabs,
                 identity_group());
}

int mom_cs_prtsrf_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                       double ri_im, int32_t ngen, const int32_t *planes, const double *x, double theta, double phi,
                       double psi, int32_t nq, const double *qw, const double *qnodes, double *csca, double *cabs) {
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  return cs_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, theta, phi, psi, nq, qw, qnodes, csca, cabs, G);
}

int mom_srcvec_pw_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                       double ri_im, int32_t ngen, const int32_t *planes, int64_t nsrc, const double *angles,
                       int32_t nq, const double *qw, const double *qnodes, double *q, int64_t ldq) {
  if (!ctx || !mesh || !angles || !q) return fail(MOM_ERR_ARG, "NULL argument");
  if (nsrc < 1 || nedges_total < 1 || ldq < 2 * nedges_total || mesh->max_parent >= nedges_total)
    return fail(MOM_ERR_ARG, "bad sizes");
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const int64_t ncol = nsrc * G.n;
  DevBuf dq;
  if ((rc = dq.alloc(sizeof(double2) * ldq * ncol))) return rc;
  MOM_HIP(hipMemsetAsync(dq.ptr, 0, sizeof(double2) * ldq * ncol, ctx->stream));
  if ((rc = srcvec_pw(ctx, mesh, rule, omega, cd(ri_re, ri_im), nsrc, angles, nedges_total, dq.as<double2>(), ldq,
                      G)))
    return rc;
  return copy_d2h(ctx, q, dq.ptr, sizeof(double2) * ldq * ncol);
}

static int far_fields_impl(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, const double *x, double r, int64_t ndir, const double *theta,
                           const double *phi, int32_t nq, const double *qw, const double *qnodes, double *out,
                           const Group &G) {
  if (!ctx || !mesh || !x || !theta || !phi || !out) return fail(MOM_ERR_ARG, "NULL argument");
  if (!(omega > 0) || !(r > 0)) return fail(MOM_ERR_ARG, "omega and r must be > 0");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "nedges_total too small");
  if (ndir < 1 || ndir > (1 << 24)) return fail(MOM_ERR_ARG, "number of directions out of range");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const int64_t N2 = 2 * nedges_total;
  DevBuf dx, dth, dph, dout, work;
  if ((rc = dx.alloc(sizeof(double2) * N2 * G.n)) || (rc = dth.alloc(sizeof(double) * ndir)) ||
      (rc = dph.alloc(sizeof(double) * ndir)) || (rc = dout.alloc(sizeof(double) * 8 * ndir)) ||
      (rc = work.alloc(far_work_bytes(mesh, rule.Q, G, ndir))))
    return rc;
  MOM_TRY(copy_h2d(ctx, dx.ptr, x, sizeof(double2) * N2 * G.n));
  MOM_TRY(copy_h2d(ctx, dth.ptr, theta, sizeof(double) * ndir));
  MOM_TRY(copy_h2d(ctx, dph.ptr, phi, sizeof(double) * ndir));
  MOM_TRY(far_fields_device(ctx, mesh, rule, omega, cd(ri_re, ri_im), dx.as<double2>(), nedges_total, r, ndir,
                            dth.as<double>(), dph.as<double>(), dout.as<double>(), work.ptr, G, N2));
  return copy_d2h(ctx, out, dout.ptr, sizeof(double) * 8 * ndir);
}

int mom_far_fields(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                   double ri_im, const double *x, double r, int64_t ndir, const double *theta, const double *phi,
                   int32_t nq, const double *qw, const double *qnodes, double *out) {
  return far_fields_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, r, ndir, theta, phi, nq, qw, qnodes, out,
                         identity_group());
}

int mom_far_fields_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, double r, int64_t ndir, const double *theta, const double *phi,
                           const double *v, int32_t nq, const double *qw, const double *qnodes, double *g) {
  if (!ctx || !mesh || !theta || !phi || !v || !g) return fail(MOM_ERR_ARG, "NULL argument");
  if (mesh->ctx != ctx) return fail(MOM_ERR_ARG, "mesh belongs to another context");
  if (!(omega > 0) || !(r > 0)) return fail(MOM_ERR_ARG, "omega and r must be > 0");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "nedges_total too small");
  QuadRule rule;
  MOM_TRY(make_rule(nq, qw, qnodes, &rule));
  (void)hipSetDevice(ctx->device);
  return far_fields_adjoint(ctx, mesh, rule, omega, cd(ri_re, ri_im), nedges_total, r, ndir, theta, phi, v, g);
}

static int rcs_impl(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                    double ri_im, const double *x, int32_t ntheta, int32_t nphi, int32_t filter, const double *fdir,
                    int32_t nq, const double *qw, const double *qnodes, double *rcs, const Group &G) {
  if (!ctx || !mesh || !x || !rcs || (filter && !fdir)) return fail(MOM_ERR_ARG, "NULL argument");
  if (ntheta < 2 || nphi < 2) return fail(MOM_ERR_ARG, "ntheta and nphi must be >= 2");
  const int64_t nd = (int64_t)ntheta * nphi;
  // ffields.f90:233-234: theta = pi (n-1)/(ntheta-1), phi = 2 pi (m-1)/(nphi-1), both ends included
  std::vector<double> th(nd), ph(nd), f(8 * nd);
  for (int m = 0; m < nphi; m++)
    for (int n = 0; n < ntheta; n++) {
      th[(int64_t)m * ntheta + n] = MOM_PI * (double)n / (double)(ntheta - 1);
      ph[(int64_t)m * ntheta + n] = 2.0 * MOM_PI * (double)m / (double)(nphi - 1);
    }
  int rc = far_fields_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, 1.0, nd, th.data(), ph.data(), nq, qw,
                           qnodes, f.data(), G);
  if (rc) return rc;
  for (int64_t d = 0; d < nd; d++) {
    const cd et(f[8 * d], f[8 * d + 1]), ep(f[8 * d + 2], f[8 * d + 3]);
    if (!filter) {
      rcs[d] = std::norm(et) + std::norm(ep);  // ffields.f90:238, c = 1
    } else {
      // ffields.f90:240-243: |fdir . e|^2, e = et v_theta + ep v_phi (dotc does not conjugate)
      const double t = th[d], p = ph[d];
      const double vt[3] = {std::cos(p) * std::cos(t), std::sin(p) * std::cos(t), -std::sin(t)};
      const double vp[3] = {-std::sin(p), std::cos(p), 0.0};
      cd s = 0;
      for (int c = 0; c < 3; c++) s += fdir[c] * (et * vt[c] + ep * vp[c]);
      rcs[d] = std::norm(s);
    }
  }
  return 0;
}

int mom_far_fields_noalloc(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, const double *x, double r, int64_t ndir, const double *theta,
                           const double *phi, int32_t nq, const double *qw, const double *qnodes, double *out) {
  if (!ctx || !mesh || !x || !theta || !phi || !out) return fail(MOM_ERR_ARG, "NULL argument");
  if (!(omega > 0) || !(r > 0)) return fail(MOM_ERR_ARG, "omega and r must be > 0");
  if (nedges_total < 1 || mesh->max_parent >= nedges_total) return fail(MOM_ERR_ARG, "nedges_total too small");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  return far_fields_no_alloc(ctx, mesh, rule, omega, cd(ri_re, ri_im), reinterpret_cast<const double2 *>(x),
                             nedges_total, r, ndir, theta, phi, out);
}

int mom_rcs_ext(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                double ri_im, const double *x, int32_t ntheta, int32_t nphi, int32_t filter, const double *fdir,
                int32_t nq, const double *qw, const double *qnodes, double *rcs) {
  return rcs_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, ntheta, nphi, filter, fdir, nq, qw, qnodes, rcs,
                  identity_group());
}

int mom_rcs_ext_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                     double ri_im, int32_t ngen, const int32_t *planes, const double *x, int32_t ntheta, int32_t nphi,
                     int32_t filter, const double *fdir, int32_t nq, const double *qw, const double *qnodes,
                     double *rcs) {
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  return rcs_impl(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, ntheta, nphi, filter, fdir, nq, qw, qnodes, rcs,
                  G);
}

}  // extern "C"

extern "C" int mom_sysmat_pmchwt(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms,
                                 const double *ri, int64_t nedges_total, double omega, int32_t nq,
                                 const double *qw, const double *qnodes, double *A, int64_t lda) {
  if (!ctx || !doms || !ri || !A) return fail(MOM_ERR_ARG, "NULL argument");
  const int64_t N2 = 2 * nedges_total;
  if (ndom < 1 || nedges_total < 1 || lda < N2) return fail(MOM_ERR_ARG, "bad sizes");
  (void)hipSetDevice(ctx->device);
  DevBuf dA;
  int rc = dA.alloc(sizeof(double2) * N2 * N2);
  if (rc) return rc;
  if ((rc = mom::fill_system(ctx, ndom, doms, ri, nedges_total, omega, nq, qw, qnodes,
                             dA.as<double2>(), N2)))
    return rc;
  return mom::copy_d2h_2d(ctx, A, lda * sizeof(double2), dA.ptr, N2 * sizeof(double2), N2 * sizeof(double2), N2);
}

namespace mom {
// sysmat_pmchwt (sysmat.f90:18-105) into a device matrix Z (2N x 2N, ldz):
// zeroed, then each domain adds D+H to the J-J block, Zsi (D+H) to the M-M
// block, K to the M-J block and -K to the J-M block in global numbering.
int fill_system(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *ri,
                int64_t nedges_total, double omega, int32_t nq, const double *qw,
                const double *qnodes, double2 *Z, int64_t ldz) {
  if (!(omega > 0)) return fail(MOM_ERR_ARG, "omega must be > 0");
  QuadRule rule;
  int rc = make_rule(nq, qw, qnodes, &rule);
  if (rc) return rc;
  for (int d = 0; d < ndom; d++) {
    if (!doms[d]) return fail(MOM_ERR_ARG, "NULL domain mesh");
    if (doms[d]->ctx != ctx) return fail(MOM_ERR_ARG, "domain mesh belongs to another context");
    if (doms[d]->max_parent >= nedges_total)
      return fail(MOM_ERR_ARG, "domain edge parent index >= nedges_total");
  }
  const int64_t N2 = 2 * nedges_total;
  MOM_HIP(hipMemset2DAsync(Z, ldz * sizeof(double2), 0, N2 * sizeof(double2), N2, ctx->stream));
  for (int d = 0; d < ndom; d++) {
    const cd r(ri[2 * d], ri[2 * d + 1]);
    FillTarget t;
    t.Z = Z;
    t.ld = ldz;
    t.Noff = nedges_total;
    t.mode = 0;
    const cd zsi = (r * r) / (kEta0 * kEta0);
    t.zsi = make_double2(zsi.real(), zsi.imag());
    if ((rc = fill_domain(ctx, doms[d], rule, omega, r, t))) return rc;
  }
  return 0;
}
}  // namespace mom
