/*
 * mom_oracle.c -- CPU restatement of the for90-mom2 PMCHWT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (for90-mom2_amd/) never links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the reference algorithm function by
 * function.  Each routine cites the reference file:line it follows.  The
 * operation order of the reference is kept where it is cheap to do so, so
 * that this restatement and the reference Fortran agree to rounding level.
 *
 * Pinning (what checks this file against something other than itself):
 *   - constants, quadrature tables: bit values of the float32-rounded
 *     literals (tests/test_oracle_tables.py);
 *   - mesh numbering/orientation, geometry, computeD/H/K, sysmat_pmchwt,
 *     srcvec, cs_prtsrf, far fields, and the second-harmonic chain
 *     (nlsurf_coef, sysmat_pmchwt_nls, the SH solution and its rcs): the
 *     reference Fortran built by oracle/build_ref.sh (golden vectors under
 *     tests/golden/);
 *   - the int.f90 closed forms: brute-force adaptive quadrature of the same
 *     integrals (tests/test_oracle_singular.py);
 *   - the whole chain: Mie series for spheres (tests/test_oracle_mie.py) and
 *     operator symmetry D = D^T, H = H^T, K = K^T.
 *
 * Indices are 0-based here; the reference is 1-based.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ---- constants.f90:8-16 ------------------------------------------------ */
/* pi is a default-REAL literal in the reference: its value is the float32
 * rounding of 3.141592653589 widened to double. */
#define ORC_PI ((double)3.141592653589f)
static const double orc_eps0 = 8.8541878176e-12;
static const double orc_c0 = 2.997924580003e8;
static const double orc_eta0 = 3.767303134622e2;
static double orc_mu0(void) { return 4 * ORC_PI * 1e-7; }

/* ---- quad.f90:115-171 tri_gl13 (float32 literals widened) ------------- */
static const float gl13_w_f[13] = {
    -0.149570044467670f, 0.175615257433204f, 0.175615257433204f,
    0.175615257433204f,  0.053347235608839f, 0.053347235608839f,
    0.053347235608839f,  0.077113760890257f, 0.077113760890257f,
    0.077113760890257f,  0.077113760890257f, 0.077113760890257f,
    0.077113760890257f};
/* quadFactors13(13,3), column-major: column 1, column 2, column 3 */
static const float gl13_n_f[39] = {
    0.333333333333333f, 0.479308067841923f, 0.260345966079038f,
    0.260345966079038f, 0.869739794195568f, 0.065130102902216f,
    0.065130102902216f, 0.638444188569809f, 0.638444188569809f,
    0.312865496004875f, 0.312865496004875f, 0.048690315425316f,
    0.048690315425316f,
    0.333333333333333f, 0.260345966079038f, 0.260345966079038f,
    0.479308067841923f, 0.065130102902216f, 0.065130102902216f,
    0.869739794195568f, 0.048690315425316f, 0.312865496004875f,
    0.048690315425316f, 0.638444188569809f, 0.312865496004875f,
    0.638444188569809f,
    0.333333333333333f, 0.260345966079038f, 0.479308067841923f,
    0.260345966079038f, 0.065130102902216f, 0.869739794195568f,
    0.065130102902216f, 0.312865496004875f, 0.048690315425316f,
    0.638444188569809f, 0.048690315425316f, 0.638444188569809f,
    0.312865496004875f};

/* Export the gl13 rule: w[13], nodes[13*3] column-major (quad.f90:305-310). */
void orc_quad_gl13(double *w, double *nodes) {
  for (int i = 0; i < 13; i++) w[i] = (double)gl13_w_f[i];
  for (int i = 0; i < 39; i++) nodes[i] = (double)gl13_n_f[i];
}

void orc_constants(double *out) {
  out[0] = ORC_PI;
  out[1] = orc_eps0;
  out[2] = orc_mu0();
  out[3] = orc_c0;
  out[4] = orc_eta0;
  out[5] = ORC_PI / 180.0;
}

typedef struct {
  int Q;
  double w[32];
  double nodes[96]; /* (Q,3) column-major */
} oquad;

/* ---- mesh.f90 types (27-83) -------------------------------------------- */
typedef struct {
  int nnodes, nfaces, nedges;
  double *p;           /* 3*nnodes */
  int *node_parent;    /* nnodes: master node index (submesh only) */
  int *fnodes;         /* 3*nfaces */
  int *fedges;         /* 3*nfaces */
  int *fid;            /* nfaces */
  int *fparent;        /* nfaces: master face (submesh only) */
  double *fn, *fcp;    /* 3*nfaces */
  double *fs, *fm;     /* 9*nfaces: s(:,k) at [9f+3k+c] */
  double *farea;       /* nfaces */
  int *enodes, *ebn, *efaces; /* 2*nedges each */
  double *elen;
  int *eparent;        /* nedges */
  int *nchild;         /* master: children per edge */
  int *child;          /* master: up to 8 (dom, ledge) pairs per edge */
  double avelen;
} omesh;

typedef struct {
  omesh master;
  int ndom;
  omesh *dom;
} oproblem;

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    exit(1);
  }
  return p;
}

static double dotr(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* linalg.f90:54-59 normr = SQRT(SUM(v*v)) */
static double normr(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}
/* linalg.f90:61-68 crossr */
static void crossr(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* mesh.f90:1018-1027 indexrot3 (1-based) */
static int indexrot3(int n) {
  int idx = n;
  if (idx > 3) idx = idx - ((idx - 1) / 3) * 3;
  return idx;
}

static int cmp_pairs(int a0, int a1, int b0, int b1) {
  return (a0 == b0 && a1 == b1) || (a0 == b1 && a1 == b0);
}

/* mesh.f90:138-153 average_edge_length */
static double average_edge_length(const omesh *m) {
  double res = 0.0;
  for (int n = 0; n < m->nedges; n++) {
    const double *a = m->p + 3 * m->enodes[2 * n];
    const double *b = m->p + 3 * m->enodes[2 * n + 1];
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    res = res + normr(d);
  }
  return res / m->nedges;
}

/* mesh.f90:1265-1322 compute_basis_data */
static void compute_basis_data(omesh *m) {
  m->fn = xcalloc(3 * m->nfaces, sizeof(double));
  m->fcp = xcalloc(3 * m->nfaces, sizeof(double));
  m->fs = xcalloc(9 * m->nfaces, sizeof(double));
  m->fm = xcalloc(9 * m->nfaces, sizeof(double));
  m->farea = xcalloc(m->nfaces, sizeof(double));
  m->elen = xcalloc(m->nedges, sizeof(double));
  for (int n = 0; n < m->nfaces; n++) {
    const double *p1 = m->p + 3 * m->fnodes[3 * n];
    const double *p2 = m->p + 3 * m->fnodes[3 * n + 1];
    const double *p3 = m->p + 3 * m->fnodes[3 * n + 2];
    double a[3], c[3], b[3], p21[3], p32[3], p13[3], mp13[3], t[3];
    for (int k = 0; k < 3; k++) {
      a[k] = p2[k] - p1[k];
      c[k] = p3[k] - p1[k];
    }
    crossr(a, c, b);
    double nb = normr(b);
    for (int k = 0; k < 3; k++) m->fn[3 * n + k] = b[k] / nb;
    for (int k = 0; k < 3; k++) {
      p21[k] = p2[k] - p1[k];
      p32[k] = p3[k] - p2[k];
      p13[k] = p1[k] - p3[k];
    }
    double *s = m->fs + 9 * n, *mm = m->fm + 9 * n;
    double l21 = normr(p21), l32 = normr(p32), l13 = normr(p13);
    for (int k = 0; k < 3; k++) {
      s[k] = p21[k] / l21;
      s[3 + k] = p32[k] / l32;
      s[6 + k] = p13[k] / l13;
    }
    for (int e = 0; e < 3; e++) crossr(s + 3 * e, m->fn + 3 * n, mm + 3 * e);
    for (int e = 0; e < 3; e++) {
      double l = normr(mm + 3 * e);
      for (int k = 0; k < 3; k++) mm[3 * e + k] = mm[3 * e + k] / l;
    }
    for (int k = 0; k < 3; k++) mp13[k] = -p13[k];
    crossr(mp13, p32, t);
    m->farea[n] = 0.5 * normr(t);
    for (int k = 0; k < 3; k++) m->fcp[3 * n + k] = (p1[k] + p2[k] + p3[k]) / 3.0;
  }
  for (int n = 0; n < m->nedges; n++) {
    const double *a = m->p + 3 * m->enodes[2 * n];
    const double *b = m->p + 3 * m->enodes[2 * n + 1];
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    m->elen[n] = normr(d);
  }
}

/* mesh.f90:1044-1137 build_mesh: first-appearance edge numbering.  The
 * reference scans the edge list linearly; a hash on the sorted node pair
 * gives the same first-appearance numbering. */
static void build_mesh(omesh *m, double scale) {
  for (int i = 0; i < 3 * m->nnodes; i++) m->p[i] = m->p[i] * scale;
  int cap = 3 * m->nfaces;
  int *en = xcalloc(2 * cap, sizeof(int)), *eb = xcalloc(2 * cap, sizeof(int)),
      *ef = xcalloc(2 * cap, sizeof(int));
  for (int i = 0; i < 2 * cap; i++) en[i] = eb[i] = ef[i] = -1;
  m->fedges = xcalloc(3 * m->nfaces, sizeof(int));
  /* open-addressing hash keyed by (min,max) node pair */
  size_t hs = 1;
  while (hs < (size_t)cap * 2) hs <<= 1;
  int *ht = xcalloc(hs, sizeof(int));
  for (size_t i = 0; i < hs; i++) ht[i] = -1;
  int cedge = 0;
  for (int n = 0; n < m->nfaces; n++) {
    for (int mm = 1; mm <= 3; mm++) {
      int a = m->fnodes[3 * n + mm - 1];
      int b = m->fnodes[3 * n + indexrot3(mm + 1) - 1];
      int bn = m->fnodes[3 * n + indexrot3(mm + 2) - 1];
      unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
      size_t h = (size_t)((lo * 2654435761ULL) ^ (hi * 40503ULL)) & (hs - 1);
      int found = -1;
      while (ht[h] >= 0) {
        int l = ht[h];
        if (cmp_pairs(en[2 * l], en[2 * l + 1], a, b)) {
          found = l;
          break;
        }
        h = (h + 1) & (hs - 1);
      }
      if (found >= 0) {
        /* an edge shared by more than two faces keeps the last one as face 2
         * (mesh.f90:1082-1086) */
        ef[2 * found + 1] = n;
        eb[2 * found + 1] = bn;
        m->fedges[3 * n + mm - 1] = found;
      } else {
        en[2 * cedge] = a;
        en[2 * cedge + 1] = b;
        ef[2 * cedge] = n;
        eb[2 * cedge] = bn;
        m->fedges[3 * n + mm - 1] = cedge;
        ht[h] = cedge;
        cedge++;
      }
    }
  }
  free(ht);
  m->nedges = cedge;
  m->enodes = en;
  m->ebn = eb;
  m->efaces = ef;
  m->eparent = xcalloc(cedge, sizeof(int));
  for (int i = 0; i < cedge; i++) m->eparent[i] = -1;
  m->avelen = average_edge_length(m);
  compute_basis_data(m);
}

/* mesh.f90:278-288 invert_faces */
static void invert_faces(omesh *m, int id) {
  for (int n = 0; n < m->nfaces; n++)
    if (m->fid[n] == id) {
      int t = m->fnodes[3 * n];
      m->fnodes[3 * n] = m->fnodes[3 * n + 2];
      m->fnodes[3 * n + 2] = t;
    }
}

/* mesh.f90:290-376 extract_submesh (surface part) */
static void extract_submesh(const omesh *M, const int *ids, int nids, omesh *S) {
  memset(S, 0, sizeof(*S));
  char *mask = xcalloc(M->nnodes, 1);
  int nf = 0;
  for (int n = 0; n < M->nfaces; n++) {
    int hit = 0;
    for (int i = 0; i < nids; i++) hit += (ids[i] == M->fid[n]);
    if (hit) {
      nf++;
      for (int k = 0; k < 3; k++) mask[M->fnodes[3 * n + k]] = 1;
    }
  }
  int *loc = xcalloc(M->nnodes, sizeof(int));
  int nn = 0;
  for (int n = 0; n < M->nnodes; n++) nn += mask[n];
  S->nnodes = nn;
  S->nfaces = nf;
  S->p = xcalloc(3 * nn, sizeof(double));
  S->node_parent = xcalloc(nn, sizeof(int));
  int m = 0;
  for (int n = 0; n < M->nnodes; n++) {
    loc[n] = -1;
    if (mask[n]) {
      for (int k = 0; k < 3; k++) S->p[3 * m + k] = M->p[3 * n + k];
      S->node_parent[m] = n;
      loc[n] = m;
      m++;
    }
  }
  S->fnodes = xcalloc(3 * nf, sizeof(int));
  S->fid = xcalloc(nf, sizeof(int));
  S->fparent = xcalloc(nf, sizeof(int));
  m = 0;
  for (int n = 0; n < M->nfaces; n++) {
    int hit = 0;
    for (int i = 0; i < nids; i++) hit += (ids[i] == M->fid[n]);
    if (hit) {
      for (int k = 0; k < 3; k++) S->fnodes[3 * m + k] = loc[M->fnodes[3 * n + k]];
      S->fparent[m] = n;
      S->fid[m] = M->fid[n];
      m++;
    }
  }
  free(mask);
  free(loc);
}

#define MAXCHILD 8

/* mesh.f90:378-452 submesh_edge_connectivity (non-periodic part). */
static int submesh_edge_connectivity(omesh *M, omesh *subs, int nsub) {
  M->nchild = xcalloc(M->nedges, sizeof(int));
  M->child = xcalloc(2 * MAXCHILD * (size_t)M->nedges, sizeof(int));
  /* master edge lookup by node pair */
  size_t hs = 1;
  while (hs < (size_t)M->nedges * 2) hs <<= 1;
  int *ht = xcalloc(hs, sizeof(int));
  for (size_t i = 0; i < hs; i++) ht[i] = -1;
  for (int n = 0; n < M->nedges; n++) {
    int a = M->enodes[2 * n], b = M->enodes[2 * n + 1];
    unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
    size_t h = (size_t)((lo * 2654435761ULL) ^ (hi * 40503ULL)) & (hs - 1);
    while (ht[h] >= 0) h = (h + 1) & (hs - 1);
    ht[h] = n;
  }
  /* The reference loops n (master edge) outer, m (submesh), l (local edge)
   * inner, so each master edge's children are ordered by (m, l). */
  for (int m = 0; m < nsub; m++) {
    omesh *S = subs + m;
    for (int l = 0; l < S->nedges; l++) {
      int a = S->node_parent[S->enodes[2 * l]], b = S->node_parent[S->enodes[2 * l + 1]];
      unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
      size_t h = (size_t)((lo * 2654435761ULL) ^ (hi * 40503ULL)) & (hs - 1);
      while (ht[h] >= 0) {
        int n = ht[h];
        if (cmp_pairs(M->enodes[2 * n], M->enodes[2 * n + 1], a, b)) {
          S->eparent[l] = n;
          int c = M->nchild[n]++;
          if (c < MAXCHILD) {
            M->child[2 * (MAXCHILD * (size_t)n + c)] = m;
            M->child[2 * (MAXCHILD * (size_t)n + c) + 1] = l;
          }
          break;
        }
        h = (h + 1) & (hs - 1);
      }
    }
  }
  free(ht);
  for (int n = 0; n < M->nedges; n++)
    if (M->nchild[n] < 2 || M->nchild[n] > MAXCHILD) {
      fprintf(stderr, "Submesh decomposition is inconsistent (invalid child edges)!\n");
      return -1;
    }
  return 0;
}

static void swapi(int *a, int *b) {
  int t = *a;
  *a = *b;
  *b = t;
}

/* mesh.f90:454-558 orient_basis (non-periodic part) */
static void orient_basis(omesh *M, omesh *subs) {
  for (int n = 0; n < M->nedges; n++) {
    int nchild = M->nchild[n];
    int oriented[MAXCHILD] = {0};
    oriented[0] = 1;
    int *ch = M->child + 2 * MAXCHILD * (size_t)n;
    int m, exited = 0;
    for (m = 1; m <= nchild; m++) {
      for (int l = 1; l < nchild; l++) {
        if (oriented[l]) continue;
        for (int k = 0; k < nchild; k++) {
          if (!oriented[k]) continue;
          omesh *Sl = subs + ch[2 * l], *Sk = subs + ch[2 * k];
          int el = ch[2 * l + 1], ek = ch[2 * k + 1];
          int pf1 = Sl->efaces[2 * el], nf1 = Sl->efaces[2 * el + 1];
          int pf2 = Sk->efaces[2 * ek], nf2 = Sk->efaces[2 * ek + 1];
          if (pf1 != -1) pf1 = Sl->fparent[pf1];
          if (nf1 != -1) nf1 = Sl->fparent[nf1];
          if (pf2 != -1) pf2 = Sk->fparent[pf2];
          if (nf2 != -1) nf2 = Sk->fparent[nf2];
          if (pf1 == pf2 || nf1 == nf2) {
            swapi(Sl->efaces + 2 * el, Sl->efaces + 2 * el + 1);
            swapi(Sl->ebn + 2 * el, Sl->ebn + 2 * el + 1);
          }
          if (pf1 == pf2 || nf1 == nf2 || pf1 == nf2 || nf1 == pf2) {
            oriented[l] = 1;
            break;
          }
        }
      }
      int cnt = 0;
      for (int i = 0; i < nchild; i++) cnt += oriented[i];
      if (cnt == nchild) {
        exited = 1;
        break;
      }
    }
    /* The reference tests the DO variable after the loop: an EXIT in the
     * last iteration leaves m == nchild and is reported as a failure that
     * stops the orientation of the remaining edges (mesh.f90:552-555). */
    if (exited && m == nchild) {
      fprintf(stderr, "Basis orientation was unsuccessful!\n");
      break;
    }
  }
}

/* ---- gmsh 2.x loader, mesh.f90:610-736 (triangles only) ---------------- */
static int load_gmsh(const char *fn, omesh *M) {
  FILE *f = fopen(fn, "r");
  if (!f) return -1;
  memset(M, 0, sizeof(*M));
  char line[1024];
  int nf = 0, fcap = 0;
  while (fgets(line, sizeof line, f)) {
    if (!strncmp(line, "$Nodes", 6)) {
      int nn;
      if (!fgets(line, sizeof line, f)) break;
      nn = atoi(line);
      M->nnodes = nn;
      M->p = xcalloc(3 * nn, sizeof(double));
      for (int i = 0; i < nn; i++) {
        int id;
        double x, y, z;
        if (!fgets(line, sizeof line, f)) break;
        sscanf(line, "%d %lf %lf %lf", &id, &x, &y, &z);
        M->p[3 * (id - 1)] = x;
        M->p[3 * (id - 1) + 1] = y;
        M->p[3 * (id - 1) + 2] = z;
      }
    } else if (!strncmp(line, "$Elements", 9)) {
      int ne;
      if (!fgets(line, sizeof line, f)) break;
      ne = atoi(line);
      fcap = ne;
      M->fnodes = xcalloc(3 * fcap, sizeof(int));
      M->fid = xcalloc(fcap, sizeof(int));
      for (int i = 0; i < ne; i++) {
        if (!fgets(line, sizeof line, f)) break;
        int v[32], c = 0;
        char *s = line, *e;
        while (c < 32) {
          long x = strtol(s, &e, 10);
          if (e == s) break;
          v[c++] = (int)x;
          s = e;
        }
        if (c >= 3 && v[1] == 2) {
          int nt = v[2];
          M->fid[nf] = v[3];
          M->fnodes[3 * nf] = v[3 + nt] - 1;
          M->fnodes[3 * nf + 1] = v[4 + nt] - 1;
          M->fnodes[3 * nf + 2] = v[5 + nt] - 1;
          nf++;
        }
      }
    }
  }
  fclose(f);
  M->nfaces = nf;
  return 0;
}

/* Build the master mesh and the domain meshes the way the `mesh`, `sdom`
 * and `fdom` commands do (interface.f90:357-366, 44-105, 107-123).
 * dom_ids holds, for each domain, nsurf[d] signed surface ids. */
void *orc_problem_load(const char *mshfile, double scale, int ndom, const int *nsurf,
                       const int *dom_ids) {
  oproblem *P = xcalloc(1, sizeof(oproblem));
  if (load_gmsh(mshfile, &P->master)) {
    free(P);
    return NULL;
  }
  build_mesh(&P->master, scale);
  P->ndom = ndom;
  P->dom = xcalloc(ndom, sizeof(omesh));
  const int *ids = dom_ids;
  for (int d = 0; d < ndom; d++) {
    int ab[64];
    for (int i = 0; i < nsurf[d]; i++) ab[i] = abs(ids[i]);
    extract_submesh(&P->master, ab, nsurf[d], P->dom + d);
    for (int i = 0; i < nsurf[d]; i++)
      if (ids[i] < 0) invert_faces(P->dom + d, -ids[i]);
    build_mesh(P->dom + d, 1.0);
    ids += nsurf[d];
  }
  if (submesh_edge_connectivity(&P->master, P->dom, ndom)) return NULL;
  orient_basis(&P->master, P->dom);
  return P;
}

static omesh *getmesh(void *h, int dom) {
  oproblem *P = h;
  return dom < 0 ? &P->master : P->dom + dom;
}

void orc_mesh_sizes(void *h, int dom, int *out) {
  omesh *m = getmesh(h, dom);
  out[0] = m->nnodes;
  out[1] = m->nfaces;
  out[2] = m->nedges;
}

double orc_mesh_avelen(void *h, int dom) { return getmesh(h, dom)->avelen; }

/* Copy out mesh arrays (0-based indices). */
void orc_mesh_export(void *h, int dom, double *p, int *fnodes, int *fedges, double *fn,
                     double *fcp, double *fs, double *fm, double *farea, int *enodes,
                     int *ebn, int *efaces, double *elen, int *eparent) {
  omesh *m = getmesh(h, dom);
  memcpy(p, m->p, 3 * m->nnodes * sizeof(double));
  memcpy(fnodes, m->fnodes, 3 * m->nfaces * sizeof(int));
  memcpy(fedges, m->fedges, 3 * m->nfaces * sizeof(int));
  memcpy(fn, m->fn, 3 * m->nfaces * sizeof(double));
  memcpy(fcp, m->fcp, 3 * m->nfaces * sizeof(double));
  memcpy(fs, m->fs, 9 * m->nfaces * sizeof(double));
  memcpy(fm, m->fm, 9 * m->nfaces * sizeof(double));
  memcpy(farea, m->farea, m->nfaces * sizeof(double));
  memcpy(enodes, m->enodes, 2 * m->nedges * sizeof(int));
  memcpy(ebn, m->ebn, 2 * m->nedges * sizeof(int));
  memcpy(efaces, m->efaces, 2 * m->nedges * sizeof(int));
  memcpy(elen, m->elen, m->nedges * sizeof(double));
  memcpy(eparent, m->eparent, m->nedges * sizeof(int));
}

/* ---- mesh.f90:1326-1369 face sign / bnodes ----------------------------- */
static double face_sign(const omesh *m, int f, int e) {
  int ge = m->fedges[3 * f + e];
  if (f == m->efaces[2 * ge]) return 1.0;
  if (f == m->efaces[2 * ge + 1]) return -1.0;
  fprintf(stderr, "Error evaluating mesh face sign!\n");
  return 0.0;
}
static const double *face_bnode(const omesh *m, int f, int e) {
  int ge = m->fedges[3 * f + e];
  if (m->efaces[2 * ge] == f) return m->p + 3 * m->ebn[2 * ge];
  return m->p + 3 * m->ebn[2 * ge + 1];
}

/* quad.f90:352-367 quad_tri_points */
static void quad_tri_points(const oquad *qd, const omesh *m, int f, double *res) {
  const double *p1 = m->p + 3 * m->fnodes[3 * f];
  const double *p2 = m->p + 3 * m->fnodes[3 * f + 1];
  const double *p3 = m->p + 3 * m->fnodes[3 * f + 2];
  for (int n = 0; n < qd->Q; n++) {
    double a = qd->nodes[2 * qd->Q + n], b = qd->nodes[n], c = qd->nodes[qd->Q + n];
    for (int k = 0; k < 3; k++) res[3 * n + k] = a * p1[k] + b * p2[k] + c * p3[k];
  }
}

/* rwgf.f90:54-78 vrwg: f = L/(2A)(r - v+) on the + face, L/(2A)(v- - r) on - */
static void vrwg(const double *r, int Q, const omesh *m, int f, int e, double *res) {
  int ge = m->fedges[3 * f + e];
  double L = m->elen[ge], A = m->farea[f];
  double s = face_sign(m, f, e);
  double c = L / (2.0 * A);
  if (s > 0) {
    const double *v = m->p + 3 * m->ebn[2 * ge];
    for (int t = 0; t < Q; t++)
      for (int k = 0; k < 3; k++) res[3 * t + k] = c * (r[3 * t + k] - v[k]);
  } else {
    const double *v = m->p + 3 * m->ebn[2 * ge + 1];
    for (int t = 0; t < Q; t++)
      for (int k = 0; k < 3; k++) res[3 * t + k] = c * (v[k] - r[3 * t + k]);
  }
}

/* rwgf.f90:34-52 rwgDiv = sign*L/A */
static double rwgDiv(const omesh *m, int f, int e) {
  int ge = m->fedges[3 * f + e];
  return face_sign(m, f, e) * m->elen[ge] / m->farea[f];
}

/* ---- green.f90 ---------------------------------------------------------- */
static cplx cexpi(cplx z) { return cexp(z); }

/* green.f90:110-121 vGf; 40-55 vGs */
static void vG(const double *r0, const double *rp, int Q, cplx k, int smooth, cplx *res) {
  const double oofpi = 1.0 / (4.0 * ORC_PI);
  for (int t = 0; t < Q; t++) {
    double dx = r0[0] - rp[3 * t], dy = r0[1] - rp[3 * t + 1], dz = r0[2] - rp[3 * t + 2];
    double R = sqrt(dx * dx + dy * dy + dz * dz);
    if (!smooth) {
      res[t] = oofpi * cexpi(I * k * R) / R;
    } else if (R == 0) {
      res[t] = k * (I / (4.0 * ORC_PI));
    } else {
      res[t] = oofpi * ((cexpi(I * k * R) - 1.0) / R + 0.5 * (k * k) * R);
    }
  }
}

/* green.f90:135-151 vgradGf; 74-97 vgradGs */
static void vgradG(const double *r0, const double *rp, int Q, cplx k, int smooth, cplx *res) {
  const double oofpi = 1.0 / (4.0 * ORC_PI);
  for (int t = 0; t < Q; t++) {
    double d[3] = {r0[0] - rp[3 * t], r0[1] - rp[3 * t + 1], r0[2] - rp[3 * t + 2]};
    double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    cplx ph;
    if (!smooth) {
      double RR = R * R;
      ph = oofpi * cexpi(I * k * R) * (1.0 / (R * RR) - I * k / RR);
    } else if (R == 0) {
      ph = 0;
    } else {
      double RR = R * R;
      ph = oofpi * (cexpi(I * k * R) * (1.0 / (R * RR) - I * k / RR) - 1.0 / (R * RR) -
                    (k * k) / (2.0 * R));
    }
    for (int c = 0; c < 3; c++) res[3 * t + c] = ph * d[c];
  }
}

/* ---- int.f90: analytic integrals (Hanninen 2006) ----------------------- */
static void edge_pts(const omesh *m, int f, int e, const double **p1, const double **p2) {
  int a = m->fnodes[3 * f + e], b = m->fnodes[3 * f + (e + 1) % 3];
  *p1 = m->p + 3 * a;
  *p2 = m->p + 3 * b;
}

/* int.f90:15-53 intLm1 */
static double intLm1(const double *r, const omesh *m, int f, int e) {
  const double *p1, *p2;
  edge_pts(m, f, e, &p1, &p2);
  const double *s = m->fs + 9 * f + 3 * e;
  double d2[3] = {r[0] - p2[0], r[1] - p2[1], r[2] - p2[2]};
  double d1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  double Rp = normr(d2), Rm = normr(d1);
  double q2[3] = {p2[0] - r[0], p2[1] - r[1], p2[2] - r[2]};
  double q1[3] = {p1[0] - r[0], p1[1] - r[1], p1[2] - r[2]};
  double sp = dotr(q2, s), sm = dotr(q1, s);
  double div1 = Rp - sp, div2 = Rm + sm;
  if (fabs(div1) > fabs(div2))
    return log((Rm - sm) / div1);
  return log((Rp + sp) / div2);
}

/* int.f90:55-96 intL1 */
static double intL1(const double *r, const omesh *m, int f, int e) {
  const double *p1, *p2;
  edge_pts(m, f, e, &p1, &p2);
  double p[3], rp[3];
  for (int k = 0; k < 3; k++) p[k] = 0.5 * (p1[k] + p2[k]);
  for (int k = 0; k < 3; k++) rp[k] = r[k] - p[k];
  double h = dotr(m->fn + 3 * f, rp);
  double t = dotr(m->fm + 9 * f + 3 * e, rp);
  double R02 = t * t + h * h;
  const double *s = m->fs + 9 * f + 3 * e;
  double d2[3] = {r[0] - p2[0], r[1] - p2[1], r[2] - p2[2]};
  double d1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  double Rp = normr(d2), Rm = normr(d1);
  double q2[3] = {p2[0] - r[0], p2[1] - r[1], p2[2] - r[2]};
  double q1[3] = {p1[0] - r[0], p1[1] - r[1], p1[2] - r[2]};
  double sp = dotr(q2, s), sm = dotr(q1, s);
  if (R02 == 0) return 0.5 * (sp * Rp - sm * Rm);
  return 0.5 * R02 * intLm1(r, m, f, e) + 0.5 * (sp * Rp - sm * Rm);
}

/* int.f90:98-135 intL3 */
static double intL3(const double *r, const omesh *m, int f, int e) {
  const double *p1, *p2;
  edge_pts(m, f, e, &p1, &p2);
  double p[3], rp[3];
  for (int k = 0; k < 3; k++) p[k] = 0.5 * (p1[k] + p2[k]);
  for (int k = 0; k < 3; k++) rp[k] = r[k] - p[k];
  double h = dotr(m->fn + 3 * f, rp);
  double t = dotr(m->fm + 9 * f + 3 * e, rp);
  double R02 = t * t + h * h;
  const double *s = m->fs + 9 * f + 3 * e;
  double d2[3] = {r[0] - p2[0], r[1] - p2[1], r[2] - p2[2]};
  double d1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  double Rp = normr(d2), Rm = normr(d1);
  double q2[3] = {p2[0] - r[0], p2[1] - r[1], p2[2] - r[2]};
  double q1[3] = {p1[0] - r[0], p1[1] - r[1], p1[2] - r[2]};
  double sp = dotr(q2, s), sm = dotr(q1, s);
  return 0.75 * R02 * intL1(r, m, f, e) + 0.25 * (sp * (Rp * Rp * Rp) - sm * (Rm * Rm * Rm));
}

/* int.f90:137-162 intSm3h: signed solid angle */
static double intSm3h(const double *r, const omesh *m, int f) {
  const double *p1 = m->p + 3 * m->fnodes[3 * f];
  const double *p2 = m->p + 3 * m->fnodes[3 * f + 1];
  const double *p3 = m->p + 3 * m->fnodes[3 * f + 2];
  double a1[3], a2[3], a3[3], d[3], c[3];
  for (int k = 0; k < 3; k++) d[k] = p1[k] - r[k];
  double l = normr(d);
  for (int k = 0; k < 3; k++) a1[k] = d[k] / l;
  for (int k = 0; k < 3; k++) d[k] = p2[k] - r[k];
  l = normr(d);
  for (int k = 0; k < 3; k++) a2[k] = d[k] / l;
  for (int k = 0; k < 3; k++) d[k] = p3[k] - r[k];
  l = normr(d);
  for (int k = 0; k < 3; k++) a3[k] = d[k] / l;
  double x = 1.0 + dotr(a1, a2) + dotr(a1, a3) + dotr(a2, a3);
  crossr(a2, a3, c);
  double y = fabs(dotr(a1, c));
  double om = 2.0 * atan2(y, x);
  double rp1[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  if (dotr(m->fn + 3 * f, rp1) > 0.0) return fabs(om);
  return -fabs(om);
}

static void face_t(const double *r, const omesh *m, int f, double *t, double *h) {
  for (int e = 0; e < 3; e++) {
    const double *pe = m->p + 3 * m->fnodes[3 * f + e];
    double d[3] = {r[0] - pe[0], r[1] - pe[1], r[2] - pe[2]};
    t[e] = dotr(m->fm + 9 * f + 3 * e, d);
  }
  const double *p1 = m->p + 3 * m->fnodes[3 * f];
  double d[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  *h = dotr(m->fn + 3 * f, d);
}

/* int.f90:181-226 intSm1 */
static double intSm1(const double *r, const omesh *m, int f) {
  double t[3], h;
  face_t(r, m, f, t, &h);
  if (h == 0) {
    double e1 = t[0] == 0 ? 0 : intLm1(r, m, f, 0);
    double e2 = t[1] == 0 ? 0 : intLm1(r, m, f, 1);
    double e3 = t[2] == 0 ? 0 : intLm1(r, m, f, 2);
    return -(t[0] * e1 + t[1] * e2 + t[2] * e3);
  }
  return -h * intSm3h(r, m, f) -
         (t[0] * intLm1(r, m, f, 0) + t[1] * intLm1(r, m, f, 1) + t[2] * intLm1(r, m, f, 2));
}

/* int.f90:228-246 intS1 */
static double intS1(const double *r, const omesh *m, int f) {
  double t[3], h;
  face_t(r, m, f, t, &h);
  return (h * h) / 3.0 * intSm1(r, m, f) -
         1.0 / 3.0 * (t[0] * intL1(r, m, f, 0) + t[1] * intL1(r, m, f, 1) + t[2] * intL1(r, m, f, 2));
}

/* int.f90:248-260 intK11, 262-274 intK1m1 */
static double intK11(const double *r, const omesh *m, int f, int e) {
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  return face_sign(m, f, e) * L / A * intS1(r, m, f);
}
static double intK1m1(const double *r, const omesh *m, int f, int e) {
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  return face_sign(m, f, e) * L / A * intSm1(r, m, f);
}

/* int.f90:276-301 intK21 */
static void intK21(const double *r, const omesh *m, int f, int e, double *res) {
  const double *p1 = m->p + 3 * m->fnodes[3 * f];
  double sign = face_sign(m, f, e);
  const double *p = face_bnode(m, f, e);
  double d[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  const double *n = m->fn + 3 * f, *mm = m->fm + 9 * f;
  double h = dotr(n, d);
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  double rho[3];
  for (int k = 0; k < 3; k++) rho[k] = r[k] - h * n[k];
  double l1 = intL3(r, m, f, 0), l2 = intL3(r, m, f, 1), l3 = intL3(r, m, f, 2);
  double s1 = intS1(r, m, f);
  double c = sign * L / (2.0 * A);
  for (int k = 0; k < 3; k++)
    res[k] = c * (1.0 / 3.0 * (mm[k] * l1 + mm[3 + k] * l2 + mm[6 + k] * l3) + (rho[k] - p[k]) * s1);
}

/* int.f90:303-328 intK2m1 */
static void intK2m1(const double *r, const omesh *m, int f, int e, double *res) {
  const double *p1 = m->p + 3 * m->fnodes[3 * f];
  double sign = face_sign(m, f, e);
  const double *p = face_bnode(m, f, e);
  double d[3] = {r[0] - p1[0], r[1] - p1[1], r[2] - p1[2]};
  const double *n = m->fn + 3 * f, *mm = m->fm + 9 * f;
  double h = dotr(n, d);
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  double rho[3];
  for (int k = 0; k < 3; k++) rho[k] = r[k] - h * n[k];
  double l1 = intL1(r, m, f, 0), l2 = intL1(r, m, f, 1), l3 = intL1(r, m, f, 2);
  double sm1 = intSm1(r, m, f);
  double c = sign * L / (2.0 * A);
  for (int k = 0; k < 3; k++)
    res[k] = c * (mm[k] * l1 + mm[3 + k] * l2 + mm[6 + k] * l3 + (rho[k] - p[k]) * sm1);
}

/* int.f90:330-354 intK31 */
static void intK31(const double *r, const omesh *m, int f, double *res) {
  const double *p = m->p + 3 * m->fnodes[3 * f];
  const double *n = m->fn + 3 * f, *mm = m->fm + 9 * f;
  double d[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]};
  double h = dotr(n, d);
  double l1 = intL1(r, m, f, 0), l2 = intL1(r, m, f, 1), l3 = intL1(r, m, f, 2);
  if (h == 0) {
    for (int k = 0; k < 3; k++) res[k] = mm[k] * l1 + mm[3 + k] * l2 + mm[6 + k] * l3;
  } else {
    double sm1 = intSm1(r, m, f);
    for (int k = 0; k < 3; k++)
      res[k] = mm[k] * l1 + mm[3 + k] * l2 + mm[6 + k] * l3 - h * n[k] * sm1;
  }
}

/* int.f90:356-374 intK3m1 */
static void intK3m1(const double *r, const omesh *m, int f, double *res) {
  const double *n = m->fn + 3 * f, *mm = m->fm + 9 * f;
  double l1 = intLm1(r, m, f, 0), l2 = intLm1(r, m, f, 1), l3 = intLm1(r, m, f, 2);
  double om = intSm3h(r, m, f);
  for (int k = 0; k < 3; k++) res[k] = mm[k] * l1 + mm[3 + k] * l2 + mm[6 + k] * l3 + n[k] * om;
}

/* int.f90:376-390 intK41, 392-406 intK4m1 */
static void intK41(const double *r, const omesh *m, int f, int e, double *res) {
  double sign = face_sign(m, f, e);
  const double *p = face_bnode(m, f, e);
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  double k31[3], d[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]}, c[3];
  intK31(r, m, f, k31);
  crossr(d, k31, c);
  double s = -sign * L / (2.0 * A);
  for (int k = 0; k < 3; k++) res[k] = s * c[k];
}
static void intK4m1(const double *r, const omesh *m, int f, int e, double *res) {
  double sign = face_sign(m, f, e);
  const double *p = face_bnode(m, f, e);
  double L = m->elen[m->fedges[3 * f + e]], A = m->farea[f];
  double k3[3], d[3] = {r[0] - p[0], r[1] - p[1], r[2] - p[2]}, c[3];
  intK3m1(r, m, f, k3);
  crossr(d, k3, c);
  double s = -sign * L / (2.0 * A);
  for (int k = 0; k < 3; k++) res[k] = s * c[k];
}

/* Export of the scalar/vector closed forms for the brute-force checks. */
double orc_int_scalar(void *h, int dom, int which, int f, int e, const double *r) {
  omesh *m = getmesh(h, dom);
  switch (which) {
  case 0: return intLm1(r, m, f, e);
  case 1: return intL1(r, m, f, e);
  case 2: return intL3(r, m, f, e);
  case 3: return intSm3h(r, m, f);
  case 4: return intSm1(r, m, f);
  case 5: return intS1(r, m, f);
  case 6: return intK11(r, m, f, e);
  case 7: return intK1m1(r, m, f, e);
  }
  return NAN;
}
void orc_int_vector(void *h, int dom, int which, int f, int e, const double *r, double *res) {
  omesh *m = getmesh(h, dom);
  switch (which) {
  case 0: intK21(r, m, f, e, res); break;
  case 1: intK2m1(r, m, f, e, res); break;
  case 2: intK31(r, m, f, res); break;
  case 3: intK3m1(r, m, f, res); break;
  case 4: intK41(r, m, f, e, res); break;
  case 5: intK4m1(r, m, f, e, res); break;
  }
}

/* ---- srcint.f90 (non-periodic, identity group) ------------------------- */
/* srcint.f90:57-118 intK1 (+ intK1singular 21-53) */
static void intK1(const double *r, const omesh *m, int f, cplx k, int near, const oquad *qd,
                  cplx *res) {
  double qpn[96];
  cplx gv[32];
  double An = m->farea[f];
  quad_tri_points(qd, m, f, qpn);
  vG(r, qpn, qd->Q, k, near, gv);
  cplx s = 0;
  for (int t = 0; t < qd->Q; t++) s += qd->w[t] * gv[t];
  cplx tmp = An * s;
  for (int e = 0; e < 3; e++) {
    double divfn = rwgDiv(m, f, e);
    cplx aint = 0;
    if (near)
      aint = 1.0 / (4.0 * ORC_PI) * intK1m1(r, m, f, e) - (k * k) / (8.0 * ORC_PI) * intK11(r, m, f, e);
    res[e] = tmp * divfn + aint;
  }
}

/* srcint.f90:158-224 intK2 (+ intK2singular 121-154); res[3*edge + comp] */
static void intK2(const double *r, const omesh *m, int f, cplx k, int near, const oquad *qd,
                  cplx *res) {
  double qpn[96], fv[96];
  cplx gv[32];
  double An = m->farea[f];
  quad_tri_points(qd, m, f, qpn);
  vG(r, qpn, qd->Q, k, near, gv);
  for (int t = 0; t < qd->Q; t++) gv[t] = gv[t] * qd->w[t] * An;
  for (int e = 0; e < 3; e++) {
    vrwg(qpn, qd->Q, m, f, e, fv);
    cplx aint[3] = {0, 0, 0};
    if (near) {
      double a[3], b[3];
      intK2m1(r, m, f, e, a);
      intK21(r, m, f, e, b);
      for (int c = 0; c < 3; c++)
        aint[c] = 1.0 / (4.0 * ORC_PI) * a[c] - (k * k) / (8.0 * ORC_PI) * b[c];
    }
    for (int c = 0; c < 3; c++) {
      cplx s = 0;
      for (int t = 0; t < qd->Q; t++) s += fv[3 * t + c] * gv[t];
      res[3 * e + c] = s + aint[c];
    }
  }
}

/* srcint.f90:370-448 intK4 (+ intK4singular 330-366) */
static void intK4(const double *r, const omesh *m, int f, cplx k, int tf, int near,
                  const oquad *qd, cplx *res) {
  if (f == tf) {
    for (int i = 0; i < 9; i++) res[i] = 0;
    return;
  }
  double qpn[96], fv[96];
  cplx ggv[96];
  double An = m->farea[f];
  quad_tri_points(qd, m, f, qpn);
  vgradG(r, qpn, qd->Q, k, near, ggv);
  for (int e = 0; e < 3; e++) {
    vrwg(qpn, qd->Q, m, f, e, fv);
    cplx aint[3] = {0, 0, 0};
    if (near) {
      double a[3], b[3];
      intK4m1(r, m, f, e, a);
      intK41(r, m, f, e, b);
      for (int c = 0; c < 3; c++)
        aint[c] = (1.0 / (4.0 * ORC_PI) * a[c] - (k * k) / (8.0 * ORC_PI) * b[c]) * 1.0;
    }
    for (int c = 0; c < 3; c++) {
      int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cplx s = 0;
      for (int t = 0; t < qd->Q; t++) {
        cplx tmp = ggv[3 * t + c1] * fv[3 * t + c2] - ggv[3 * t + c2] * fv[3 * t + c1];
        s += tmp * qd->w[t];
      }
      res[3 * e + c] = s * An + aint[c];
    }
  }
}

/* sysmat.f90:233-260 near_faces (non-periodic, identity group) */
static int near_faces(const omesh *m, int n, int mm) {
  double threshold = m->avelen * 3;
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = m->fcp[3 * mm + k] - m->fcp[3 * n + k];
  double distsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return distsq < threshold * threshold;
}

/* sysmat.f90:272-347 computeD, 359-434 computeH, 446-519 computeK.
 * op: 0 = D, 1 = H, 2 = K.  M is E x E column-major, overwritten. */
static void compute_op_rows(const omesh *m, int op, double omega, cplx ri, const oquad *qd, cplx *M,
                            int m0, int m1);
static void compute_op(const omesh *m, int op, double omega, cplx ri, const oquad *qd, cplx *M) {
  memset(M, 0, sizeof(cplx) * (size_t)m->nedges * m->nedges);
  compute_op_rows(m, op, omega, ri, qd, M, 0, m->nfaces);
}

/* computeD/H/K restricted to test faces m0 <= m < m1 (the reference's outer
 * loop, sysmat.f90:303/390/473), adding into M. */
static void compute_op_rows(const omesh *m, int op, double omega, cplx ri, const oquad *qd, cplx *M,
                            int m0, int m1) {
  int E = m->nedges, F = m->nfaces, Q = qd->Q;
  cplx k = ri * omega / orc_c0;
  cplx c1 = -I * omega * orc_mu0();
  cplx c2 = -1.0 / (I * omega * orc_eps0 * (ri * ri));
  int per = (op == 1) ? 3 : 9;
  cplx *intaux = xcalloc((size_t)per * Q * F, sizeof(cplx));
  for (int mm = m0; mm < m1; mm++) {
    double qpm[96], fmv[3][96], fmDiv[3];
    quad_tri_points(qd, m, mm, qpm);
    double Am = m->farea[mm];
    for (int p = 0; p < 3; p++) {
      vrwg(qpm, Q, m, mm, p, fmv[p]);
      fmDiv[p] = rwgDiv(m, mm, p);
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < F; n++) {
      int near = near_faces(m, n, mm);
      for (int r = 0; r < Q; r++) {
        cplx *dst = intaux + ((size_t)n * Q + r) * per;
        if (op == 0) intK2(qpm + 3 * r, m, n, k, near, qd, dst);
        else if (op == 1) intK1(qpm + 3 * r, m, n, k, near, qd, dst);
        else intK4(qpm + 3 * r, m, n, k, mm, near, qd, dst);
      }
    }
    for (int n = 0; n < F; n++) {
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
          cplx int1 = 0;
          for (int r = 0; r < Q; r++) {
            const cplx *a = intaux + ((size_t)n * Q + r) * per;
            if (op == 1) {
              int1 += qd->w[r] * a[q];
            } else {
              const double *fv = fmv[p] + 3 * r;
              cplx d = fv[0] * a[3 * q] + fv[1] * a[3 * q + 1] + fv[2] * a[3 * q + 2];
              int1 += qd->w[r] * d;
            }
          }
          if (op == 1) int1 = int1 * Am * fmDiv[p];
          else int1 = int1 * Am;
          int i1 = m->fedges[3 * mm + p], i2 = m->fedges[3 * n + q];
          size_t idx = (size_t)i2 * E + i1;
          if (op == 0) M[idx] += c1 * int1;
          else if (op == 1) M[idx] += c2 * int1;
          else M[idx] += int1;
        }
    }
  }
  free(intaux);
}

static void mkquad(oquad *qd, int Q, const double *w, const double *nodes) {
  qd->Q = Q;
  for (int i = 0; i < Q; i++) qd->w[i] = w[i];
  for (int i = 0; i < 3 * Q; i++) qd->nodes[i] = nodes[i];
}

void orc_compute_op(void *h, int dom, int op, double omega, double ri_re, double ri_im, int Q,
                    const double *w, const double *nodes, double *Mout) {
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  compute_op(getmesh(h, dom), op, omega, ri_re + I * ri_im, &qd, (cplx *)Mout);
}

/* sysmat.f90:18-105 sysmat_pmchwt (identity group, non-periodic).
 * ri: ndom complex values (interleaved); A: 2E x 2E column-major. */
void orc_sysmat(void *h, double omega, const double *ri, int Q, const double *w,
                const double *nodes, double *Aout) {
  oproblem *P = h;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int N = P->master.nedges;
  size_t N2 = 2 * (size_t)N;
  cplx *A = (cplx *)Aout;
  memset(A, 0, sizeof(cplx) * N2 * N2);
  for (int d = 0; d < P->ndom; d++) {
    omesh *m = P->dom + d;
    cplx r = ri[2 * d] + I * ri[2 * d + 1];
    cplx Zsi = (r * r) / (orc_eta0 * orc_eta0);
    int E = m->nedges;
    int *ind = m->eparent;
    cplx *M = xcalloc((size_t)E * E, sizeof(cplx));
    double detj = 1.0;
    for (int op = 0; op < 3; op++) {
      compute_op(m, op, omega, r, &qd, M);
      for (int j = 0; j < E; j++)
        for (int i = 0; i < E; i++) {
          cplx v = M[(size_t)j * E + i];
          size_t I0 = ind[i], J0 = ind[j];
          if (op < 2) {
            A[J0 * N2 + I0] += v;
            A[(J0 + N) * N2 + I0 + N] += detj * Zsi * v;
          } else {
            A[J0 * N2 + I0 + N] += v;
            A[(J0 + N) * N2 + I0] -= detj * v;
          }
        }
    }
    free(M);
  }
}

/* aux.f90:245-261 get_dir / get_pol */
static void get_dir(double th, double ph, double *d) {
  d[0] = sin(th) * cos(ph);
  d[1] = sin(th) * sin(ph);
  d[2] = cos(th);
}
static void get_pol(double th, double ph, double ps, double *p) {
  p[0] = cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph);
  p[1] = -cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph);
  p[2] = sin(ps) * sin(th);
}

/* source.f90:84-101 pw_fields (with pw_efield 63-69, pw_hfield 71-80) */
static void pw_fields(double th, double ph, double ps, double omega, cplx ri, const double *pt,
                      cplx *ef, cplx *hf) {
  double dir[3], pol[3];
  get_dir(th, ph, dir);
  get_pol(th, ph, ps, pol);
  cplx k = ri * omega / orc_c0;
  cplx ph0 = cexpi(I * k * dotr(dir, pt));
  for (int c = 0; c < 3; c++) ef[c] = pol[c] * ph0;
  cplx f = k / (omega * orc_mu0());
  hf[0] = f * (dir[1] * ef[2] - dir[2] * ef[1]);
  hf[1] = f * (dir[2] * ef[0] - dir[0] * ef[2]);
  hf[2] = f * (dir[0] * ef[1] - dir[1] * ef[0]);
}

/* source.f90:136-247 srcvec for a plane wave on domain `dom` (identity
 * group).  Angles in radians.  q has 2*Ntot complex entries. */
void orc_srcvec(void *h, int dom, double omega, double ri_re, double ri_im, double th, double ph,
                double ps, int Q, const double *w, const double *nodes, double *qout) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int Nt = P->master.nedges;
  cplx *q = (cplx *)qout, ri = ri_re + I * ri_im;
  memset(q, 0, sizeof(cplx) * 2 * (size_t)Nt);
  for (int f = 0; f < m->nfaces; f++) {
    double qpm[96];
    quad_tri_points(&qd, m, f, qpm);
    double Am = m->farea[f];
    cplx ef[32][3], hf[32][3];
    for (int r = 0; r < Q; r++) pw_fields(th, ph, ps, omega, ri, qpm + 3 * r, ef[r], hf[r]);
    for (int p = 0; p < 3; p++) {
      cplx Ie = 0, Ih = 0;
      double fv[96];
      vrwg(qpm, Q, m, f, p, fv);
      for (int r = 0; r < Q; r++) {
        const double *fm = fv + 3 * r;
        Ie += qd.w[r] * (fm[0] * ef[r][0] + fm[1] * ef[r][1] + fm[2] * ef[r][2]);
        Ih += qd.w[r] * (fm[0] * hf[r][0] + fm[1] * hf[r][1] + fm[2] * hf[r][2]);
      }
      int idx = m->eparent[m->fedges[3 * f + p]];
      q[idx] += Ie * Am;
      q[idx + Nt] += Ih * Am;
    }
  }
}

static void crossc(const cplx *a, const cplx *b, cplx *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* cs.f90:14-104 cs_prtsrf (identity group, plane-wave source). */
void orc_cs(void *h, int dom, double omega, double ri_re, double ri_im, const double *xin,
            double th, double ph, double ps, int Q, const double *w, const double *nodes,
            double *out) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int Nt = P->master.nedges;
  const cplx *x = (const cplx *)xin;
  cplx ri = ri_re + I * ri_im;
  double c = 2.0 / (orc_eps0 * orc_c0 * creal(ri));
  double csca = 0, cabs = 0;
  for (int n = 0; n < m->nfaces; n++) {
    double qpn[96];
    quad_tri_points(&qd, m, n, qpn);
    double An = m->farea[n];
    const double *nor = m->fn + 3 * n;
    double int1 = 0, int2 = 0;
    for (int t = 0; t < Q; t++) {
      cplx Et[3] = {0, 0, 0}, Ht[3] = {0, 0, 0};
      for (int p = 0; p < 3; p++) {
        double fv[3], fn[3];
        vrwg(qpn + 3 * t, 1, m, n, p, fv);
        crossr(nor, fv, fn);
        int idx = m->eparent[m->fedges[3 * n + p]];
        for (int k = 0; k < 3; k++) {
          Ht[k] = Ht[k] - x[idx] * fn[k];
          Et[k] = Et[k] + x[idx + Nt] * fn[k];
        }
      }
      cplx einc[3], hinc[3], cn[3] = {nor[0], nor[1], nor[2]}, tmp[3], Eti[3], Hti[3];
      pw_fields(th, ph, ps, omega, ri, qpn + 3 * t, einc, hinc);
      crossc(cn, hinc, tmp);
      crossc(tmp, cn, Hti);
      crossc(cn, einc, tmp);
      crossc(tmp, cn, Eti);
      cplx Hts[3], Ets[3], cH[3], cr[3];
      for (int k = 0; k < 3; k++) {
        Hts[k] = Ht[k] - Hti[k];
        Ets[k] = Et[k] - Eti[k];
      }
      for (int k = 0; k < 3; k++) cH[k] = conj(Ht[k]);
      crossc(Et, cH, cr);
      double Sn = 0.5 * creal(cr[0] * nor[0] + cr[1] * nor[1] + cr[2] * nor[2]);
      int1 += qd.w[t] * Sn;
      for (int k = 0; k < 3; k++) cH[k] = conj(Hts[k]);
      crossc(Ets, cH, cr);
      Sn = 0.5 * creal(cr[0] * nor[0] + cr[1] * nor[1] + cr[2] * nor[2]);
      int2 += qd.w[t] * Sn;
    }
    cabs += int1 * An;
    csca += int2 * An;
  }
  out[0] = c * csca;
  out[1] = -c * cabs;
}

/* ffields.f90:16-122 far_fields at r = 1 (identity group) and 203-249
 * rcs_ext without polarisation filter.  out: ntheta x nphi column-major. */
void orc_rcs(void *h, int dom, double omega, double ri_re, double ri_im, const double *xin,
             int ntheta, int nphi, int Q, const double *w, const double *nodes, double *out) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int Nt = P->master.nedges;
  const cplx *x = (const cplx *)xin;
  cplx ri = ri_re + I * ri_im;
  cplx k = ri * omega / orc_c0;
#pragma omp parallel for schedule(static) collapse(2)
  for (int a = 0; a < ntheta; a++)
    for (int b = 0; b < nphi; b++) {
      double theta = ORC_PI * (double)a / (double)(ntheta - 1);
      double phi = 2.0 * ORC_PI * (double)b / (double)(nphi - 1);
      double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
      cplx Nt_ = 0, Np = 0, Lt = 0, Lp = 0;
      for (int n = 0; n < m->nfaces; n++) {
        double An = m->farea[n], qpn[96];
        quad_tri_points(&qd, m, n, qpn);
        for (int q = 0; q < 3; q++) {
          int idx = m->eparent[m->fedges[3 * n + q]];
          cplx q1 = 0, q2 = 0, q3 = 0, q4 = 0;
          for (int t = 0; t < Q; t++) {
            const double *rp = qpn + 3 * t;
            double jg[3];
            vrwg(rp, 1, m, n, q, jg);
            cplx jn[3], mn[3];
            for (int c = 0; c < 3; c++) {
              jn[c] = jg[c] * x[idx];
              mn[c] = jg[c] * x[idx + Nt];
            }
            cplx phasor = cexpi(-I * k * (rp[0] * st * cp + rp[1] * st * sp + rp[2] * ct));
            q1 += qd.w[t] * (ct * cp * jn[0] + ct * sp * jn[1] - st * jn[2]) * phasor;
            q2 += qd.w[t] * (-sp * jn[0] + cp * jn[1]) * phasor;
            q3 += qd.w[t] * (ct * cp * mn[0] + ct * sp * mn[1] - st * mn[2]) * phasor;
            q4 += qd.w[t] * (-sp * mn[0] + cp * mn[1]) * phasor;
          }
          Nt_ += q1 * An;
          Np += q2 * An;
          Lt += q3 * An;
          Lp += q4 * An;
        }
      }
      double r = 1.0;
      cplx prefix = I * k * cexpi(I * k * r) / (4.0 * ORC_PI * r);
      cplx et = prefix * (Lp + orc_eta0 / ri * Nt_);
      cplx ep = -prefix * (Lt - orc_eta0 / ri * Np);
      out[(size_t)b * ntheta + a] = cabs(et) * cabs(et) + cabs(ep) * cabs(ep);
    }
}

/* CPU timing sample of the fill: the three operators of every domain for
 * test faces [m0, m1) of each domain mesh; returns seconds of wall time. */
#include <omp.h>
double orc_fill_rows_seconds(void *h, double omega, const double *ri, int Q, const double *w,
                             const double *nodes, int m0, int m1) {
  oproblem *P = h;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  double t0 = omp_get_wtime();
  for (int d = 0; d < P->ndom; d++) {
    omesh *m = P->dom + d;
    cplx r = ri[2 * d] + I * ri[2 * d + 1];
    cplx *M = xcalloc((size_t)m->nedges * m->nedges, sizeof(cplx));
    for (int op = 0; op < 3; op++) compute_op_rows(m, op, omega, r, &qd, M, m0, m1 < m->nfaces ? m1 : m->nfaces);
    free(M);
  }
  return omp_get_wtime() - t0;
}

/* ---- second harmonic, surface chi2 (nlsurf.f90), identity group --------- */

/* nlsurf.f90:21-49 efield: surface E of the fundamental solution x at r on
 * face f, from the normal component div J/(i omega eps) and n x M */
static void nls_efield(const omesh *m, int Nt, double omega, cplx ri, const cplx *x, int f, const double *r,
                       cplx *e) {
  cplx eps = (ri * ri) * orc_eps0;
  cplx divj = 0, ms[3] = {0, 0, 0};
  for (int n = 0; n < 3; n++) {
    int ed = m->eparent[m->fedges[3 * f + n]];
    double fv[3];
    divj = divj + x[ed] * rwgDiv(m, f, n);
    vrwg(r, 1, m, f, n, fv);
    for (int k = 0; k < 3; k++) ms[k] = ms[k] + x[Nt + ed] * fv[k];
  }
  cplx en = divj / ((I * omega) * eps);
  const double *nor = m->fn + 3 * f;
  cplx cn[3] = {nor[0], nor[1], nor[2]}, t[3];
  crossc(cn, ms, t);
  for (int k = 0; k < 3; k++) e[k] = t[k] + nor[k] * en;
}

/* nlsurf.f90:74-103 Pnls_invmap and 106-126 Pnls_frag with the identity
 * action (J = I, field factor 1): P = 2 eps0 chi_ttn e_t e_n
 * + eps0 (chi_nnn e_n^2 + chi_ntt e_t.e_t) n, dotc without conjugation */
static void nls_pnls(const omesh *m, int Nt, double omega, cplx ri, const cplx *x, int f, const cplx *chi,
                     const double *r, cplx *P) {
  cplx e[3], et[3];
  nls_efield(m, Nt, omega, ri, x, f, r, e);
  const double *nor = m->fn + 3 * f;
  cplx en = nor[0] * e[0] + nor[1] * e[1] + nor[2] * e[2];
  for (int k = 0; k < 3; k++) et[k] = e[k] - en * nor[k];
  cplx Pn = orc_eps0 * (chi[0] * (en * en) + chi[1] * (et[0] * et[0] + et[1] * et[1] + et[2] * et[2]));
  for (int k = 0; k < 3; k++) P[k] = 2.0 * orc_eps0 * chi[2] * et[k] * en + Pn * nor[k];
}

/* nlsurf_coef (nlsurf.f90:137-217): the projections before the first Gram
 * solve.  Outputs (local edge numbering, complex interleaved):
 *   coef[E]     sum_faces A <div f_m, -P_n>/epsp        (rhs 1 of the solve)
 *   srcJ[E]     -i 2omega sum_faces A <f_m, P_t>         (rhs 2)
 *   srcvec[2E]  (0.5 coef, 0.5 i 2omega sum_faces A <f_m, P_t x n>) */
void orc_nls_stage1(void *h, int dom, double omega, double ri_re, double ri_im, const double *epsp_in,
                    const double *xin, const double *chi_in, int Q, const double *w, const double *nodes,
                    double *coef_out, double *srcJ_out, double *srcvec_out) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int Nt = P->master.nedges, E = m->nedges;
  const cplx *x = (const cplx *)xin, *epsp = (const cplx *)epsp_in, *chi = (const cplx *)chi_in;
  cplx ri = ri_re + I * ri_im;
  double omegash = 2.0 * omega;
  cplx *coef = (cplx *)coef_out, *srcJ = (cplx *)srcJ_out, *sv = (cplx *)srcvec_out;
  for (int i = 0; i < E; i++) coef[i] = srcJ[i] = sv[i] = sv[i + E] = 0;
  for (int f = 0; f < m->nfaces; f++) {
    double qp[96];
    cplx Pn[32], Pt[32][3];
    quad_tri_points(&qd, m, f, qp);
    double A = m->farea[f];
    const double *nor = m->fn + 3 * f;
    cplx cn[3] = {nor[0], nor[1], nor[2]};
    for (int r = 0; r < Q; r++) {
      cplx Pv[3];
      nls_pnls(m, Nt, omega, ri, x, f, chi, qp + 3 * r, Pv);
      Pn[r] = nor[0] * Pv[0] + nor[1] * Pv[1] + nor[2] * Pv[2];
      for (int k = 0; k < 3; k++) Pt[r][k] = Pv[k] - Pn[r] * nor[k];
    }
    for (int q = 0; q < 3; q++) {
      double fmDiv = rwgDiv(m, f, q);
      cplx int1 = 0, int2 = 0, int3 = 0;
      for (int r = 0; r < Q; r++) {
        double fm[3];
        cplx c[3];
        vrwg(qp + 3 * r, 1, m, f, q, fm);
        crossc(Pt[r], cn, c);
        int1 = int1 - qd.w[r] * fmDiv * Pn[r];
        int2 = int2 + qd.w[r] * (fm[0] * c[0] + fm[1] * c[1] + fm[2] * c[2]);
        int3 = int3 + qd.w[r] * (fm[0] * Pt[r][0] + fm[1] * Pt[r][1] + fm[2] * Pt[r][2]);
      }
      int idx = m->fedges[3 * f + q];
      coef[idx] = coef[idx] + int1 * A / epsp[f];
      srcJ[idx] = srcJ[idx] + int3 * A;
      sv[idx + E] = sv[idx + E] + int2 * A;
    }
  }
  for (int i = 0; i < E; i++) {
    srcJ[i] = -srcJ[i] * I * omegash;
    sv[i] = 0.5 * coef[i];
    sv[i + E] = 0.5 * I * omegash * sv[i + E];
  }
}

/* rwgf.f90:148-221 rwg_moments: Gram matrix F[m,n] = <f_m, f_n> (E x E,
 * column-major, real), the off-diagonal entry on the first common face */
void orc_rwg_moments(void *h, int dom, int Q, const double *w, const double *nodes, double *F) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int E = m->nedges;
  memset(F, 0, sizeof(double) * (size_t)E * E);
  for (int a = 0; a < E; a++) {
    for (int s = 0; s < 2; s++) {
      int f = m->efaces[2 * a + s];
      if (f == -1) continue;
      double qp[96], fm[3], in = 0;
      quad_tri_points(&qd, m, f, qp);
      int q = m->fedges[3 * f] == a ? 0 : (m->fedges[3 * f + 1] == a ? 1 : 2);
      for (int r = 0; r < Q; r++) {
        vrwg(qp + 3 * r, 1, m, f, q, fm);
        in = in + qd.w[r] * dotr(fm, fm);
      }
      in = in * m->farea[f];
      F[a + (size_t)a * E] = F[a + (size_t)a * E] + in;
    }
    for (int b = a + 1; b < E; b++) {
      int fa0 = m->efaces[2 * a], fa1 = m->efaces[2 * a + 1], fb0 = m->efaces[2 * b], fb1 = m->efaces[2 * b + 1];
      int f;
      if (fa0 != -1 && (fa0 == fb0 || fa0 == fb1))
        f = fa0;
      else if (fa1 != -1 && (fa1 == fb0 || fa1 == fb1))
        f = fa1;
      else
        continue;
      int q = m->fedges[3 * f] == a ? 0 : (m->fedges[3 * f + 1] == a ? 1 : 2);
      int p = m->fedges[3 * f] == b ? 0 : (m->fedges[3 * f + 1] == b ? 1 : 2);
      double qp[96], fm[3], fn[3], in = 0;
      quad_tri_points(&qd, m, f, qp);
      for (int r = 0; r < Q; r++) {
        vrwg(qp + 3 * r, 1, m, f, q, fm);
        vrwg(qp + 3 * r, 1, m, f, p, fn);
        in = in + qd.w[r] * dotr(fm, fn);
      }
      in = in * m->farea[f];
      F[a + (size_t)b * E] = in;
      F[b + (size_t)a * E] = in;
    }
  }
}

/* nlsurf_coef (nlsurf.f90:278-299): srcM[m] = sum_faces A <f_m, n x g>,
 * g = sum_q coef[q] f_q (rwg_exp, rwgf.f90:278-298), coef from the first
 * Gram solve; srcM is the rhs of the second solve */
void orc_nls_stage2(void *h, int dom, int Q, const double *w, const double *nodes, const double *coef_in,
                    double *srcM_out) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  oquad qd;
  mkquad(&qd, Q, w, nodes);
  int E = m->nedges;
  const cplx *coef = (const cplx *)coef_in;
  cplx *srcM = (cplx *)srcM_out;
  for (int i = 0; i < E; i++) srcM[i] = 0;
  for (int f = 0; f < m->nfaces; f++) {
    double qp[96];
    quad_tri_points(&qd, m, f, qp);
    double A = m->farea[f];
    const double *nor = m->fn + 3 * f;
    cplx cn[3] = {nor[0], nor[1], nor[2]};
    for (int q = 0; q < 3; q++) {
      cplx int1 = 0;
      for (int r = 0; r < Q; r++) {
        double fm[3];
        cplx g[3] = {0, 0, 0}, c[3];
        for (int t = 0; t < 3; t++) {
          double ft[3];
          vrwg(qp + 3 * r, 1, m, f, t, ft);
          for (int k = 0; k < 3; k++) g[k] = g[k] + ft[k] * coef[m->fedges[3 * f + t]];
        }
        vrwg(qp + 3 * r, 1, m, f, q, fm);
        crossc(cn, g, c);
        int1 = int1 + qd.w[r] * (fm[0] * c[0] + fm[1] * c[1] + fm[2] * c[2]);
      }
      int idx = m->fedges[3 * f + q];
      srcM[idx] = srcM[idx] + int1 * A;
    }
  }
}

/* solver.f90:393-433 determine_epsp: per face of domain dom, eps0 ri^2 of
 * the first other domain holding the same master face (ri: 2*ndom doubles,
 * the fundamental indices).  Returns -1 if a face has no such domain. */
int orc_epsp(void *h, int dom, const double *ri, double *out) {
  oproblem *P = h;
  omesh *m = P->dom + dom;
  cplx *e = (cplx *)out;
  for (int f = 0; f < m->nfaces; f++) {
    int found = 0;
    for (int d = 0; d < P->ndom && !found; d++) {
      if (d == dom) continue;
      omesh *o = P->dom + d;
      for (int l = 0; l < o->nfaces; l++)
        if (o->fparent[l] == m->fparent[f]) {
          cplx r = ri[2 * d] + I * ri[2 * d + 1];
          e[f] = (r * r) * orc_eps0;
          found = 1;
          break;
        }
    }
    if (!found) return -1;
  }
  return 0;
}

void orc_problem_free(void *h) {
  /* process-lifetime test helper: memory is reclaimed at exit */
  (void)h;
}
