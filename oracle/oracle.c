/*
 * oracle.c -- FP64 CPU restatement of FEN ZI's per-step force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libfenzi's HIP path.  Nothing in
 * the product links, loads or calls this file.
 *
 * It follows the reference's functional forms (citations are to the reference
 * source tree, src/...) and deliberately departs from the reference defects
 * listed in the design notes: exact switched-LJ gradient (not the formula at
 * nonbonded_forces.cu:204-218), closed-form B-splines (not the 1001-sample
 * table of defs.in.h:172,188-192), unique-cell stencils, unique 1-4 pairs,
 * FP64 SHAKE with a relative-residual criterion.
 *
 * Units: AKMA (A, kcal/mol, amu, time unit = 1/TU ps).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/defs.in.h:156-158 */
#define OR_RC 0.001987
#define OR_CC 332.0716
#define OR_TU 20.4548
#define OR_PI 3.14159265358979323846
#define TWO_POW_64 18446744073709551616.0
#define TWO_POW_M64 (1.0 / 18446744073709551616.0)

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* System                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  int natom, ncons;
  int atom[8];
  int c1[12], c2[12];
  double d2[12];
} or_cluster;

struct or_sys {
  or_topo t;
  or_params p;
  /* exclusion sets as sorted unique (i<j) keys i*N+j */
  int64_t n_excl, n_ew, n_14;
  uint64_t *excl, *ew, *p14;
  /* per-atom CSR of all exclusions (both directions), sorted */
  int64_t *xoff;
  int32_t *xlist;
  int n_clusters;
  or_cluster *clusters;
  uint8_t *in_cluster;  /* [N]: 1 for an atom of a constraint cluster */
  int n_cons_total;
  /* fixed atoms: mobile[i] = 0 holds atom i; NULL = all mobile */
  uint8_t *mobile;
  int n_fixed;
  /* COM restraint */
  int com_on;
  int32_t *com_grp;
  double com_d[3], com_fc, com_m[2];
};

/* inverse mass of the integrator (0 for a fixed atom) */
static inline double dyn_minv(const or_sys *s, int i) {
  return (s->mobile && !s->mobile[i]) ? 0.0 : 1.0 / s->t.mass[i];
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

static int64_t uniq_u64(uint64_t *v, int64_t n) {
  if (n == 0) return 0;
  qsort(v, (size_t)n, sizeof(uint64_t), cmp_u64);
  int64_t m = 1;
  for (int64_t i = 1; i < n; i++)
    if (v[i] != v[m - 1]) v[m++] = v[i];
  return m;
}

static uint64_t pkey(int a, int b, int n) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return (uint64_t)lo * (uint64_t)n + (uint64_t)hi;
}

static int in_set(const uint64_t *v, int64_t n, uint64_t k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (v[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo < n && v[lo] == k;
}

/* Exclusion sets, restating GenExcl_minimal_new (src/prmread.cu:2928-3062):
 * 1-2 from the bond list, 1-3 from angle ends (this union is the Ewald
 * correction list ewlist, :2988), 1-4 from dihedral ends.  The 1-4 pair list is
 * made unique and excludes pairs already in the 1-2/1-3 set (the reference
 * evaluates one 1-4 interaction per dihedral, src/bonded_forces.cu:1155). */
static void build_exclusions(or_sys *s) {
  const or_topo *t = &s->t;
  int n = t->n_atoms;
  int64_t cap = (int64_t)t->n_bonds + t->n_angles + 1;
  s->ew = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap);
  int64_t k = 0;
  for (int i = 0; i < t->n_bonds; i++) {
    int a = t->bond_atoms[2 * i], b = t->bond_atoms[2 * i + 1];
    if (a != b) s->ew[k++] = pkey(a, b, n);
  }
  for (int i = 0; i < t->n_angles; i++) {
    int a = t->angle_atoms[3 * i], c = t->angle_atoms[3 * i + 2];
    if (a != c) s->ew[k++] = pkey(a, c, n);
  }
  s->n_ew = uniq_u64(s->ew, k);
  s->p14 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(t->n_dihedrals + 1));
  k = 0;
  for (int i = 0; i < t->n_dihedrals; i++) {
    int a = t->dihedral_atoms[4 * i], d = t->dihedral_atoms[4 * i + 3];
    if (a == d) continue;
    uint64_t key = pkey(a, d, n);
    if (!in_set(s->ew, s->n_ew, key)) s->p14[k++] = key;
  }
  s->n_14 = uniq_u64(s->p14, k);
  s->n_excl = s->n_ew + s->n_14;
  s->excl = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(s->n_excl + 1));
  memcpy(s->excl, s->ew, sizeof(uint64_t) * (size_t)s->n_ew);
  memcpy(s->excl + s->n_ew, s->p14, sizeof(uint64_t) * (size_t)s->n_14);
  s->n_excl = uniq_u64(s->excl, s->n_excl);
  /* CSR */
  s->xoff = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t e = 0; e < s->n_excl; e++) {
    int a = (int)(s->excl[e] / (uint64_t)n), b = (int)(s->excl[e] % (uint64_t)n);
    s->xoff[a + 1]++; s->xoff[b + 1]++;
  }
  for (int i = 0; i < n; i++) s->xoff[i + 1] += s->xoff[i];
  s->xlist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->xoff[n] + 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  for (int i = 0; i < n; i++) fill[i] = s->xoff[i];
  for (int64_t e = 0; e < s->n_excl; e++) {
    int a = (int)(s->excl[e] / (uint64_t)n), b = (int)(s->excl[e] % (uint64_t)n);
    s->xlist[fill[a]++] = b; s->xlist[fill[b]++] = a;
  }
  free(fill);
  for (int i = 0; i < n; i++) {
    int32_t *L = s->xlist + s->xoff[i];
    int64_t m = s->xoff[i + 1] - s->xoff[i];
    for (int64_t x = 1; x < m; x++) { /* insertion sort; lists are short */
      int32_t v = L[x]; int64_t y = x - 1;
      while (y >= 0 && L[y] > v) { L[y + 1] = L[y]; y--; }
      L[y + 1] = v;
    }
  }
}

static int excluded(const or_sys *s, int a, int b) {
  const int32_t *L = s->xlist + s->xoff[a];
  int64_t lo = 0, hi = s->xoff[a + 1] - s->xoff[a];
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (L[mid] < b) lo = mid + 1; else hi = mid;
  }
  return lo < s->xoff[a + 1] - s->xoff[a] && L[lo] == b;
}

/* Constraint clusters (connected components of the constraint graph), the
 * unit the reference's SHAKE solves together (src/prmread.cu:1192-1303). */
static int uf_find(int *p, int x) { while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; } return x; }

static void build_clusters(or_sys *s) {
  const or_topo *t = &s->t;
  int n = t->n_atoms, nc = t->n_constraints;
  free(s->clusters);
  free(s->in_cluster);
  s->n_clusters = 0; s->clusters = NULL; s->n_cons_total = 0;
  s->in_cluster = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
  if (nc == 0) return;
  int *par = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) par[i] = i;
  for (int c = 0; c < nc; c++) {
    int a0 = t->constraint_atoms[2 * c], b0 = t->constraint_atoms[2 * c + 1];
    if (s->mobile && !s->mobile[a0] && !s->mobile[b0]) continue;
    int a = uf_find(par, a0), b = uf_find(par, b0);
    if (a != b) par[a] = b;
  }
  int *cid = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) cid[i] = -1;
  s->clusters = (or_cluster *)calloc((size_t)nc, sizeof(or_cluster));
  for (int c = 0; c < nc; c++) {
    int a = t->constraint_atoms[2 * c], b = t->constraint_atoms[2 * c + 1];
    if (s->mobile && !s->mobile[a] && !s->mobile[b]) continue; /* both fixed */
    s->n_cons_total++;
    int r = uf_find(par, a);
    if (cid[r] < 0) cid[r] = s->n_clusters++;
    or_cluster *cl = &s->clusters[cid[r]];
    int la = -1, lb = -1;
    for (int k = 0; k < cl->natom; k++) { if (cl->atom[k] == a) la = k; if (cl->atom[k] == b) lb = k; }
    if (la < 0) { la = cl->natom; cl->atom[cl->natom++] = a; }
    if (lb < 0) { lb = cl->natom; cl->atom[cl->natom++] = b; }
    s->in_cluster[a] = s->in_cluster[b] = 1;
    if (cl->natom > 8 || cl->ncons >= 12) { fprintf(stderr, "oracle: constraint cluster too large\n"); abort(); }
    cl->c1[cl->ncons] = la; cl->c2[cl->ncons] = lb;
    cl->d2[cl->ncons] = t->constraint_length[c] * t->constraint_length[c];
    cl->ncons++;
  }
  free(par); free(cid);
}

or_sys *or_create(const or_topo *t, const or_params *p) {
  or_sys *s = (or_sys *)calloc(1, sizeof(or_sys));
  s->t = *t; s->p = *p;
  build_exclusions(s);
  build_clusters(s);
  return s;
}

void or_destroy(or_sys *s) {
  if (!s) return;
  free(s->excl); free(s->ew); free(s->p14); free(s->xoff); free(s->xlist); free(s->clusters); free(s->in_cluster);
  free(s->mobile); free(s->com_grp);
  free(s);
}

int or_set_fixed(or_sys *s, const uint8_t *fixed) {
  int n = s->t.n_atoms, nf = 0;
  if (fixed)
    for (int i = 0; i < n; i++) nf += fixed[i] ? 1 : 0;
  if (nf == n) return -1;
  free(s->mobile);
  s->mobile = NULL;
  s->n_fixed = nf;
  if (nf > 0) {
    s->mobile = (uint8_t *)malloc((size_t)n);
    for (int i = 0; i < n; i++) s->mobile[i] = fixed[i] ? 0 : 1;
  }
  build_clusters(s);
  return 0;
}

int or_set_com_restraint(or_sys *s, const int32_t *group, const double dist[3], double fc) {
  int n = s->t.n_atoms;
  s->com_on = 0;
  if (!group || fc == 0.0) return 0;
  free(s->com_grp);
  s->com_grp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  s->com_m[0] = s->com_m[1] = 0.0;
  for (int i = 0; i < n; i++) {
    s->com_grp[i] = group[i];
    if (group[i] == 0 || group[i] == 1) s->com_m[group[i]] += s->t.mass[i];
  }
  if (!(s->com_m[0] > 0 && s->com_m[1] > 0)) return -1;
  for (int d = 0; d < 3; d++) s->com_d[d] = dist[d];
  s->com_fc = fc;
  s->com_on = 1;
  return 0;
}

/* group COMs from positions in [-L/2, L/2) (the reference's wrap,
 * src/cell_based_kernels.cu:738-763), force fc (d + c0 - c1) on group 1 */
static void com_restraint(const or_sys *s, const uint64_t *u, double *f) {
  int n = s->t.n_atoms;
  double sum[2][3] = {{0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; i++) {
    int g = s->com_grp[i];
    if (g < 0) continue;
    for (int d = 0; d < 3; d++) sum[g][d] += s->t.mass[i] * ((double)(int64_t)u[3 * i + d] * s->p.box[d] * TWO_POW_M64);
  }
  double dr[3];
  for (int d = 0; d < 3; d++) {
    double x = s->com_d[d] + (sum[0][d] / s->com_m[0] - sum[1][d] / s->com_m[1]);
    x -= s->p.box[d] * rint(x / s->p.box[d]);
    dr[d] = s->com_fc * x;
  }
  for (int i = 0; i < n; i++)
    if (s->com_grp[i] == 1)
      for (int d = 0; d < 3; d++) f[3 * i + d] += dr[d];
}

int or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n; return 1;
#endif
}

int64_t or_n_pairs(or_sys *s, int which) { return which == 0 ? s->n_excl : which == 1 ? s->n_ew : s->n_14; }

void or_get_pairs(or_sys *s, int which, int32_t *ij) {
  const uint64_t *v = which == 0 ? s->excl : which == 1 ? s->ew : s->p14;
  int64_t m = or_n_pairs(s, which);
  uint64_t n = (uint64_t)s->t.n_atoms;
  for (int64_t e = 0; e < m; e++) { ij[2 * e] = (int32_t)(v[e] / n); ij[2 * e + 1] = (int32_t)(v[e] % n); }
}

/* ------------------------------------------------------------------------ */
/* Fixed point                                                              */
/* ------------------------------------------------------------------------ */

void or_to_fixed(const double box[3], const double *x, uint64_t *u, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double t = x[3 * i + d] / box[d];
      t -= floor(t);
      if (!(t < 1.0)) t = 0.0;
      u[3 * i + d] = (uint64_t)(t * TWO_POW_64);
    }
}

void or_from_fixed(const double box[3], const uint64_t *u, double *x, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double v = (double)u[3 * i + d] * (box[d] * TWO_POW_M64);
      if (v >= 0.5 * box[d]) v -= box[d];
      x[3 * i + d] = v;
    }
}

static inline void fx_disp(const uint64_t *u, int i, int j, const double sc[3], double d[3]) {
  /* r_j - r_i, minimum image */
  for (int k = 0; k < 3; k++) d[k] = (double)(int64_t)(u[3 * j + k] - u[3 * i + k]) * sc[k];
}

/* ------------------------------------------------------------------------ */
/* Pair interactions                                                        */
/* ------------------------------------------------------------------------ */

/* Real-space pair: erfc Coulomb + switched LJ (src/nonbonded_forces.cu:146-233,
 * switch from src/mem.cu:557-560) with the exact switched-LJ gradient.
 * Returns force/r (positive = repulsive) and the two energies. */
static inline double pair_real(double r2, double qq, double eps, double sig2, double beta,
                               double rc2, double ron2, int do_switch, double *eC, double *eLJ) {
  double r = sqrt(r2);
  double br = beta * r;
  double ec = OR_CC * qq * erfc(br) / r;
  double fC = (ec + OR_CC * qq * (2.0 * beta / sqrt(OR_PI)) * exp(-br * br)) / r2;
  double s2 = sig2 / r2, s6 = s2 * s2 * s2;
  double a = eps * s6, b = a * s6;
  double elj = b - a;
  double fLJ = (12.0 * b - 6.0 * a) / r2;
  if (do_switch && r2 > ron2) {
    if (r2 >= rc2) { fLJ = 0.0; elj = 0.0; }
    else {
      double D = (rc2 - ron2) * (rc2 - ron2) * (rc2 - ron2);
      double c1 = rc2 - r2, c2 = rc2 + 2.0 * r2 - 3.0 * ron2;
      double S = c1 * c1 * c2 / D;
      double dS = 12.0 * c1 * (r2 - ron2) / D; /* = -(dS/dr)/r */
      fLJ = fLJ * S + elj * dS;
      elj *= S;
    }
  }
  *eC = ec; *eLJ = elj;
  return fC + fLJ;
}

static void add_pair_virial(double *vir, const double d[3], double fr) {
  /* pair virial r_ij (x) F_ij = fr * d (x) d */
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) vir[3 * a + b] += fr * d[a] * d[b];
}

static void real_space(or_sys *s, const uint64_t *u, double *f, double *e, double *vir) {
  const or_topo *t = &s->t; const or_params *p = &s->p;
  int n = t->n_atoms;
  double sc[3] = {p->box[0] * TWO_POW_M64, p->box[1] * TWO_POW_M64, p->box[2] * TWO_POW_M64};
  double rc2 = p->cutoff * p->cutoff, ron2 = p->cuton * p->cuton;
  /* cell grid with edge >= cutoff; stencil restricted to unique cells */
  int nc[3];
  for (int d = 0; d < 3; d++) { nc[d] = (int)floor(p->box[d] / p->cutoff); if (nc[d] < 1) nc[d] = 1; }
  int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * (size_t)ncell), *next = (int *)malloc(sizeof(int) * (size_t)n);
  int *cellof = (int *)malloc(sizeof(int) * (size_t)n);
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int i = n - 1; i >= 0; i--) {
    int c[3];
    for (int d = 0; d < 3; d++) c[d] = (int)(((u128)u[3 * i + d] * (u128)nc[d]) >> 64);
    int id = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    cellof[i] = id; next[i] = head[id]; head[id] = i;
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  double *fth = (double *)calloc((size_t)nth * 3 * (size_t)n, sizeof(double));
  double *eth = (double *)calloc((size_t)nth * 2, sizeof(double));
  double *vth = (double *)calloc((size_t)nth * 9, sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *ft = fth + (size_t)tid * 3 * n, *vt = vth + 9 * tid;
    int ci = cellof[i];
    int c0[3] = {ci % nc[0], (ci / nc[0]) % nc[1], ci / (nc[0] * nc[1])};
    int rng[3][3], nr[3];
    for (int d = 0; d < 3; d++) {
      if (nc[d] < 3) { nr[d] = nc[d]; for (int k = 0; k < nc[d]; k++) rng[d][k] = k; }
      else { nr[d] = 3; for (int k = 0; k < 3; k++) rng[d][k] = (c0[d] + k - 1 + nc[d]) % nc[d]; }
    }
    double ei = 0, eli = 0;
    for (int a = 0; a < nr[2]; a++)
      for (int b = 0; b < nr[1]; b++)
        for (int c = 0; c < nr[0]; c++) {
          int cj = (rng[2][a] * nc[1] + rng[1][b]) * nc[0] + rng[0][c];
          for (int j = head[cj]; j >= 0; j = next[j]) {
            if (j <= i) continue;
            double d[3];
            fx_disp(u, i, j, sc, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rc2) continue;
            if (excluded(s, i, j)) continue;
            int ti = t->type[i], tj = t->type[j];
            double eC, eLJ;
            double fr = pair_real(r2, t->charge[i] * t->charge[j], t->lj_eps[ti * t->n_types + tj],
                                  t->lj_sig2[ti * t->n_types + tj], p->beta, rc2, ron2, 1, &eC, &eLJ);
            ei += eC; eli += eLJ;
            for (int k = 0; k < 3; k++) { ft[3 * i + k] -= fr * d[k]; ft[3 * j + k] += fr * d[k]; }
            add_pair_virial(vt, d, fr);
          }
        }
    eth[2 * tid] += ei; eth[2 * tid + 1] += eli;
  }
  for (int th = 0; th < nth; th++) {
    for (int64_t k = 0; k < 3 * (int64_t)n; k++) f[k] += fth[(size_t)th * 3 * n + k];
    e[OR_E_ELEC] += eth[2 * th]; e[OR_E_VDW] += eth[2 * th + 1];
    for (int k = 0; k < 9; k++) vir[k] += vth[9 * th + k];
  }
  free(fth); free(eth); free(vth); free(head); free(next); free(cellof);
}

/* Non-excluded pairs with r < rcut, over a cell grid of edge >= rcut whose
 * stencil visits each neighbouring cell once (fewer than three cells on an axis
 * visit every cell of that axis once), so each pair is tested exactly once. */
int64_t or_count_pairs_within(or_sys *s, const uint64_t *u, double rcut) {
  int n = s->t.n_atoms;
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  double r2c = rcut * rcut;
  int nc[3];
  for (int d = 0; d < 3; d++) { nc[d] = (int)floor(s->p.box[d] / rcut); if (nc[d] < 1) nc[d] = 1; }
  int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * (size_t)ncell), *next = (int *)malloc(sizeof(int) * (size_t)n);
  int *cellof = (int *)malloc(sizeof(int) * (size_t)n);
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int i = n - 1; i >= 0; i--) {
    int c[3];
    for (int d = 0; d < 3; d++) c[d] = (int)(((u128)u[3 * i + d] * (u128)nc[d]) >> 64);
    int id = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    cellof[i] = id; next[i] = head[id]; head[id] = i;
  }
  int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cnt)
  for (int i = 0; i < n; i++) {
    int ci = cellof[i];
    int c0[3] = {ci % nc[0], (ci / nc[0]) % nc[1], ci / (nc[0] * nc[1])};
    int rng[3][3], nr[3];
    for (int d = 0; d < 3; d++) {
      if (nc[d] < 3) { nr[d] = nc[d]; for (int k = 0; k < nc[d]; k++) rng[d][k] = k; }
      else { nr[d] = 3; for (int k = 0; k < 3; k++) rng[d][k] = (c0[d] + k - 1 + nc[d]) % nc[d]; }
    }
    for (int a = 0; a < nr[2]; a++)
      for (int b = 0; b < nr[1]; b++)
        for (int c = 0; c < nr[0]; c++) {
          int cj = (rng[2][a] * nc[1] + rng[1][b]) * nc[0] + rng[0][c];
          for (int j = head[cj]; j >= 0; j = next[j]) {
            if (j <= i) continue;
            double d[3];
            fx_disp(u, i, j, sc, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < r2c && !excluded(s, i, j)) cnt++;
          }
        }
  }
  free(head); free(next); free(cellof);
  return cnt;
}

/* Ewald exclusion correction for 1-2/1-3 pairs (src/mdcuda.cu:520-553) and
 * 1-4 pairs with full erfc Coulomb + 1-4 LJ (src/bonded_forces.cu:1155-1268). */
static void excl_terms(or_sys *s, const uint64_t *u, double *f, double *e, double *vir) {
  const or_topo *t = &s->t; const or_params *p = &s->p;
  int n = t->n_atoms;
  double sc[3] = {p->box[0] * TWO_POW_M64, p->box[1] * TWO_POW_M64, p->box[2] * TWO_POW_M64};
  double beta = p->beta;
  for (int64_t k = 0; k < s->n_ew; k++) {
    int i = (int)(s->ew[k] / (uint64_t)n), j = (int)(s->ew[k] % (uint64_t)n);
    double d[3];
    fx_disp(u, i, j, sc, d); /* r_j - r_i */
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
    double qq = OR_CC * t->charge[i] * t->charge[j];
    double br = beta * r, erfv = erf(br);
    double en = -qq * erfv / r;
    double fr = qq * (br * 2.0 / sqrt(OR_PI) * exp(-br * br) - erfv) / (r2 * r); /* F_i = fr*(r_i-r_j) */
    e[OR_E_EWEXCL] += en;
    for (int c = 0; c < 3; c++) { f[3 * i + c] -= fr * d[c]; f[3 * j + c] += fr * d[c]; }
    add_pair_virial(vir, d, fr);
  }
  double rc2 = p->cutoff * p->cutoff, ron2 = p->cuton * p->cuton;
  for (int64_t k = 0; k < s->n_14; k++) {
    int i = (int)(s->p14[k] / (uint64_t)n), j = (int)(s->p14[k] % (uint64_t)n);
    double d[3];
    fx_disp(u, i, j, sc, d);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    int ti = t->type[i], tj = t->type[j];
    double eC, eLJ;
    double fr = pair_real(r2, t->charge[i] * t->charge[j], t->lj14_eps[ti * t->n_types + tj],
                          t->lj14_sig2[ti * t->n_types + tj], beta, rc2, ron2, 1, &eC, &eLJ);
    e[OR_E_ELEC] += eC; e[OR_E_VDW] += eLJ;
    for (int c = 0; c < 3; c++) { f[3 * i + c] -= fr * d[c]; f[3 * j + c] += fr * d[c]; }
    add_pair_virial(vir, d, fr);
  }
}

/* ------------------------------------------------------------------------ */
/* Bonded terms (src/bonded_forces.cu:22-392)                                */
/* ------------------------------------------------------------------------ */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Dihedral angle of (a,b,c,d) (IUPAC sign) and its gradient w.r.t. the four
 * positions, in the Blondel-Karplus form (J. Comput. Chem. 17, 1132 (1996)). */
static double dihedral_grad(const double *F, const double *G, const double *H, double g[4][3]) {
  /* F = r_a - r_b, G = r_b - r_c, H = r_d - r_c */
  double A[3], B[3];
  cross3(F, G, A); cross3(H, G, B);
  double A2 = dot3(A, A), B2 = dot3(B, B), Gn = sqrt(dot3(G, G));
  double BxA[3]; cross3(B, A, BxA);
  double sinp = dot3(BxA, G) / (sqrt(A2 * B2) * Gn);
  double cosp = dot3(A, B) / sqrt(A2 * B2);
  double phi = atan2(sinp, cosp);
  double FG = dot3(F, G), HG = dot3(H, G);
  for (int k = 0; k < 3; k++) {
    g[0][k] = -Gn / A2 * A[k];
    g[3][k] = Gn / B2 * B[k];
    g[1][k] = Gn / A2 * A[k] + FG / (A2 * Gn) * A[k] - HG / (B2 * Gn) * B[k];
    g[2][k] = HG / (B2 * Gn) * B[k] - FG / (A2 * Gn) * A[k] - Gn / B2 * B[k];
  }
  return phi;
}

static void bonded(or_sys *s, const uint64_t *u, double *f, double *e, double *vir) {
  const or_topo *t = &s->t; const or_params *p = &s->p;
  double sc[3] = {p->box[0] * TWO_POW_M64, p->box[1] * TWO_POW_M64, p->box[2] * TWO_POW_M64};
  for (int i = 0; i < t->n_bonds; i++) {
    int a = t->bond_atoms[2 * i], b = t->bond_atoms[2 * i + 1];
    double d[3]; fx_disp(u, a, b, sc, d); /* r_b - r_a */
    double r = sqrt(dot3(d, d)), dr = r - t->bond_r0[i];
    e[OR_E_BOND] += t->bond_k[i] * dr * dr;
    double fr = -2.0 * t->bond_k[i] * dr / r; /* F_b = fr*d */
    for (int c = 0; c < 3; c++) { f[3 * b + c] += fr * d[c]; f[3 * a + c] -= fr * d[c]; }
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) vir[3 * x + y] += d[x] * fr * d[y];
  }
  for (int i = 0; i < t->n_angles; i++) {
    int a = t->angle_atoms[3 * i], b = t->angle_atoms[3 * i + 1], c = t->angle_atoms[3 * i + 2];
    double v1[3], v2[3];
    fx_disp(u, b, a, sc, v1); fx_disp(u, b, c, sc, v2); /* r_a - r_b, r_c - r_b */
    double l1 = dot3(v1, v1), l2 = dot3(v2, v2), q = 1.0 / sqrt(l1 * l2);
    double cs = dot3(v1, v2) * q;
    if (fabs(cs) > 0.9999) cs = copysign(0.9999, cs); /* src/bonded_forces.cu:81 */
    double sn = sqrt(1.0 - cs * cs), th = acos(cs), dth = th - t->angle_theta0[i];
    e[OR_E_ANGLE] += t->angle_k[i] * dth * dth;
    double kf = 2.0 * t->angle_k[i] * dth / sn;
    double fa[3], fc[3];
    for (int k = 0; k < 3; k++) {
      fa[k] = kf * (v2[k] * q - cs * v1[k] / l1);
      fc[k] = kf * (v1[k] * q - cs * v2[k] / l2);
      f[3 * a + k] += fa[k]; f[3 * c + k] += fc[k]; f[3 * b + k] -= fa[k] + fc[k];
    }
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) vir[3 * x + y] += v1[x] * fa[y] + v2[x] * fc[y];
    if (t->ub_k && t->ub_k[i] != 0.0) { /* Urey-Bradley 1-3 (src/bonded_forces.cu:935-991) */
      double d[3]; fx_disp(u, a, c, sc, d);
      double r = sqrt(dot3(d, d)), dr = r - t->ub_s0[i];
      e[OR_E_UB] += t->ub_k[i] * dr * dr;
      double fr = -2.0 * t->ub_k[i] * dr / r;
      for (int k = 0; k < 3; k++) { f[3 * c + k] += fr * d[k]; f[3 * a + k] -= fr * d[k]; }
      for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) vir[3 * x + y] += d[x] * fr * d[y];
    }
  }
  for (int pass = 0; pass < 2; pass++) {
    int nt = pass == 0 ? t->n_dihedrals : t->n_impropers;
    const int32_t *at = pass == 0 ? t->dihedral_atoms : t->improper_atoms;
    for (int i = 0; i < nt; i++) {
      int ia = at[4 * i], ib = at[4 * i + 1], ic = at[4 * i + 2], id = at[4 * i + 3];
      double F[3], G[3], H[3], rb[3], rc_[3], rd[3];
      fx_disp(u, ib, ia, sc, F);   /* r_a - r_b */
      fx_disp(u, ic, ib, sc, G);   /* r_b - r_c */
      fx_disp(u, ic, id, sc, H);   /* r_d - r_c */
      double g[4][3];
      double phi = dihedral_grad(F, G, H, g);
      double dEdphi;
      if (pass == 0) {
        /* E = K (1 + cos(n phi - delta)) (src/bonded_forces.cu:344-372) */
        double nn = (double)t->dihedral_n[i], dl = t->dihedral_delta[i], K = t->dihedral_k[i];
        e[OR_E_DIHED] += K * (1.0 + cos(nn * phi - dl));
        dEdphi = -K * nn * sin(nn * phi - dl);
      } else {
        /* E = K psi^2, psi0 not used (src/bonded_forces.cu:196,216) */
        double K = t->improper_k[i];
        e[OR_E_IMPR] += K * phi * phi;
        dEdphi = 2.0 * K * phi;
      }
      int idx[4] = {ia, ib, ic, id};
      double fo[4][3];
      for (int m = 0; m < 4; m++)
        for (int k = 0; k < 3; k++) { fo[m][k] = -dEdphi * g[m][k]; f[3 * idx[m] + k] += fo[m][k]; }
      /* virial with positions relative to c: r_a-r_c = F - G... r_a - r_c = (r_a-r_b)+(r_b-r_c) */
      for (int k = 0; k < 3; k++) { rb[k] = G[k]; rc_[k] = 0.0; rd[k] = H[k]; }
      double ra[3] = {F[0] + G[0], F[1] + G[1], F[2] + G[2]};
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++)
          vir[3 * x + y] += ra[x] * fo[0][y] + rb[x] * fo[1][y] + rc_[x] * fo[2][y] + rd[x] * fo[3][y];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FFT (recursive mixed radix, unnormalised)                                */
/* ------------------------------------------------------------------------ */

static void dft_rec(int n, const double *in, int istride, double *out, int sign, double *tmp) {
  /* out[k] = sum_j in[j*istride] e^{sign 2 pi i jk/n} (complex interleaved) */
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  int p = 2;
  while (n % p) p++;
  int m = n / p;
  if (p == n && n > 5) { /* prime: direct */
    for (int k = 0; k < n; k++) {
      double re = 0, im = 0;
      for (int j = 0; j < n; j++) {
        double a = sign * 2.0 * OR_PI * (double)(((long)j * k) % n) / n;
        double c = cos(a), s_ = sin(a);
        const double *x = in + 2 * (size_t)j * istride;
        re += x[0] * c - x[1] * s_; im += x[0] * s_ + x[1] * c;
      }
      out[2 * k] = re; out[2 * k + 1] = im;
    }
    return;
  }
  /* sub-transforms of the p decimated sequences into tmp */
  for (int r = 0; r < p; r++)
    dft_rec(m, in + 2 * (size_t)r * istride, istride * p, tmp + 2 * (size_t)r * m, sign, out);
  for (int k = 0; k < n; k++) {
    double re = 0, im = 0;
    for (int r = 0; r < p; r++) {
      double a = sign * 2.0 * OR_PI * (double)(((long)r * k) % n) / n;
      double c = cos(a), s_ = sin(a);
      const double *x = tmp + 2 * ((size_t)r * m + (k % m));
      re += x[0] * c - x[1] * s_; im += x[0] * s_ + x[1] * c;
    }
    out[2 * k] = re; out[2 * k + 1] = im;
  }
}

static void fft1d(int n, double *data, int stride, int sign) {
  double *in = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *out = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int j = 0; j < n; j++) { in[2 * j] = data[2 * (size_t)j * stride]; in[2 * j + 1] = data[2 * (size_t)j * stride + 1]; }
  /* dft_rec uses out as scratch for deeper levels; give it a separate buffer */
  double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  dft_rec(n, in, 1, out, sign, tmp);
  (void)scr;
  for (int j = 0; j < n; j++) { data[2 * (size_t)j * stride] = out[2 * j]; data[2 * (size_t)j * stride + 1] = out[2 * j + 1]; }
  free(in); free(out); free(tmp); free(scr);
}

void or_fft3d(const int32_t K[3], const double *in, double *out, int sign) {
  size_t tot = (size_t)K[0] * K[1] * K[2];
  if (out != in) memcpy(out, in, sizeof(double) * 2 * tot);
  /* axis 0 (fastest) */
#pragma omp parallel for
  for (long l = 0; l < (long)K[1] * K[2]; l++) fft1d(K[0], out + 2 * (size_t)l * K[0], 1, sign);
#pragma omp parallel for
  for (long l = 0; l < (long)K[0] * K[2]; l++) {
    long x = l % K[0], z = l / K[0];
    fft1d(K[1], out + 2 * ((size_t)z * K[1] * K[0] + x), K[0], sign);
  }
#pragma omp parallel for
  for (long l = 0; l < (long)K[0] * K[1]; l++) fft1d(K[2], out + 2 * (size_t)l, K[0] * K[1], sign);
}

/* ------------------------------------------------------------------------ */
/* Smooth PME, order 4 (src/pme_kernels.cu, src/cell_based_kernels.cu:1770-1947) */
/* ------------------------------------------------------------------------ */

static inline void bspline4(double w, int *k0, double th[4], double dth[4]) {
  double fl = floor(w);
  double f = w - fl;
  *k0 = (int)fl;
  double f2 = f * f, f3 = f2 * f, g = 1.0 - f;
  th[0] = g * g * g / 6.0;
  th[1] = (3.0 * f3 - 6.0 * f2 + 4.0) / 6.0;
  th[2] = (-3.0 * f3 + 3.0 * f2 + 3.0 * f + 1.0) / 6.0;
  th[3] = f3 / 6.0;
  dth[0] = -0.5 * g * g;
  dth[1] = 0.5 * (3.0 * f2 - 4.0 * f);
  dth[2] = 0.5 * (-3.0 * f2 + 2.0 * f + 1.0);
  dth[3] = 0.5 * f2;
}

static void pme(or_sys *s, const uint64_t *u, double *f, double *e, double *vir) {
  const or_topo *t = &s->t; const or_params *p = &s->p;
  int n = t->n_atoms;
  const int *K = p->K;
  size_t tot = (size_t)K[0] * K[1] * K[2];
  double *Q = (double *)calloc(2 * tot, sizeof(double));
  int *k0 = (int *)malloc(sizeof(int) * 3 * (size_t)n);
  double *th = (double *)malloc(sizeof(double) * 12 * (size_t)n), *dth = (double *)malloc(sizeof(double) * 12 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double w = (double)u[3 * i + d] * TWO_POW_M64 * K[d];
      bspline4(w, &k0[3 * i + d], th + 12 * i + 4 * d, dth + 12 * i + 4 * d);
    }
  for (int i = 0; i < n; i++) {
    double q = t->charge[i];
    if (q == 0.0) continue;
    const double *tx = th + 12 * i, *ty = tx + 4, *tz = tx + 8;
    for (int c = 0; c < 4; c++) {
      int gz = ((k0[3 * i + 2] - 3 + c) % K[2] + K[2]) % K[2];
      for (int b = 0; b < 4; b++) {
        int gy = ((k0[3 * i + 1] - 3 + b) % K[1] + K[1]) % K[1];
        for (int a = 0; a < 4; a++) {
          int gx = ((k0[3 * i] - 3 + a) % K[0] + K[0]) % K[0];
          Q[2 * (((size_t)gz * K[1] + gy) * K[0] + gx)] += q * tx[a] * ty[b] * tz[c];
        }
      }
    }
  }
  or_fft3d(K, Q, Q, -1);
  double V = p->box[0] * p->box[1] * p->box[2];
  double beta = p->beta, E = 0, W[9] = {0};
  for (int jz = 0; jz < K[2]; jz++)
    for (int jy = 0; jy < K[1]; jy++)
      for (int jx = 0; jx < K[0]; jx++) {
        size_t idx = ((size_t)jz * K[1] + jy) * K[0] + jx;
        if (idx == 0) { Q[0] = Q[1] = 0; continue; }
        int js[3] = {jx, jy, jz};
        double m[3], B = 1.0;
        for (int d = 0; d < 3; d++) {
          int jj = js[d] <= K[d] / 2 ? js[d] : js[d] - K[d];
          m[d] = jj / p->box[d];
          double c = cos(2.0 * OR_PI * js[d] / K[d]);
          B *= 9.0 / ((2.0 + c) * (2.0 + c));
        }
        double m2 = dot3(m, m);
        double G = OR_CC / (OR_PI * V) * exp(-OR_PI * OR_PI * m2 / (beta * beta)) / m2 * B;
        double S2 = Q[2 * idx] * Q[2 * idx] + Q[2 * idx + 1] * Q[2 * idx + 1];
        double Em = 0.5 * G * S2;
        E += Em;
        double fac = 2.0 * (1.0 + OR_PI * OR_PI * m2 / (beta * beta)) / m2;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) W[3 * a + b] += Em * ((a == b ? 1.0 : 0.0) - fac * m[a] * m[b]);
        Q[2 * idx] *= G; Q[2 * idx + 1] *= G;
      }
  or_fft3d(K, Q, Q, +1);
  e[OR_E_PME] += E;
  for (int k = 0; k < 9; k++) vir[k] += W[k];
#pragma omp parallel for
  for (int i = 0; i < n; i++) {
    double q = t->charge[i];
    if (q == 0.0) continue;
    const double *tx = th + 12 * i, *ty = tx + 4, *tz = tx + 8;
    const double *dx = dth + 12 * i, *dy = dx + 4, *dz = dx + 8;
    double g[3] = {0, 0, 0};
    for (int c = 0; c < 4; c++) {
      int gz = ((k0[3 * i + 2] - 3 + c) % K[2] + K[2]) % K[2];
      for (int b = 0; b < 4; b++) {
        int gy = ((k0[3 * i + 1] - 3 + b) % K[1] + K[1]) % K[1];
        for (int a = 0; a < 4; a++) {
          int gx = ((k0[3 * i] - 3 + a) % K[0] + K[0]) % K[0];
          double ph = Q[2 * (((size_t)gz * K[1] + gy) * K[0] + gx)];
          g[0] += dx[a] * ty[b] * tz[c] * ph;
          g[1] += tx[a] * dy[b] * tz[c] * ph;
          g[2] += tx[a] * ty[b] * dz[c] * ph;
        }
      }
    }
    for (int d = 0; d < 3; d++) f[3 * i + d] -= q * g[d] * K[d] / p->box[d];
  }
  free(Q); free(k0); free(th); free(dth);
}

/* Direct Ewald k-space sum (test reference for PME on small systems). */
static void ewald_direct(or_sys *s, const uint64_t *u, double *f, double *e, double *vir) {
  const or_topo *t = &s->t; const or_params *p = &s->p;
  int n = t->n_atoms;
  double beta = p->beta, V = p->box[0] * p->box[1] * p->box[2];
  double tol = p->ewald_kmax_tol > 0 ? p->ewald_kmax_tol : 1e-18;
  double mmax = beta / OR_PI * sqrt(-log(tol));
  int nmax[3];
  for (int d = 0; d < 3; d++) nmax[d] = (int)ceil(mmax * p->box[d]);
  int W0 = 2 * nmax[0] + 1, W1 = 2 * nmax[1] + 1, W2 = 2 * nmax[2] + 1;
  /* per-atom phase tables e^{2 pi i n t_d}, n in [-nmax, nmax] */
  double *ph = (double *)malloc(sizeof(double) * 2 * (size_t)n * (W0 + W1 + W2));
#pragma omp parallel for
  for (int i = 0; i < n; i++) {
    double *base = ph + 2 * (size_t)i * (W0 + W1 + W2);
    int off = 0;
    for (int d = 0; d < 3; d++) {
      int Wd = 2 * nmax[d] + 1;
      double tt = (double)u[3 * i + d] * TWO_POW_M64;
      for (int k = -nmax[d]; k <= nmax[d]; k++) {
        double a = 2.0 * OR_PI * (double)k * tt;
        base[2 * (off + k + nmax[d])] = cos(a);
        base[2 * (off + k + nmax[d]) + 1] = sin(a);
      }
      off += Wd;
    }
  }
  double A = OR_CC / (2.0 * OR_PI * V);
  double E = 0, Wv[9] = {0};
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  double *fth = (double *)calloc((size_t)nth * 3 * n, sizeof(double));
  double *eth = (double *)calloc((size_t)nth * 10, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int kz = -nmax[2]; kz <= nmax[2]; kz++) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *ft = fth + (size_t)tid * 3 * n, *et = eth + 10 * tid;
    double *re_j = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int ky = -nmax[1]; ky <= nmax[1]; ky++)
      for (int kx = -nmax[0]; kx <= nmax[0]; kx++) {
        /* half space */
        if (kx < 0 || (kx == 0 && ky < 0) || (kx == 0 && ky == 0 && kz <= 0)) continue;
        double m[3] = {kx / p->box[0], ky / p->box[1], kz / p->box[2]};
        double m2 = dot3(m, m);
        double g = exp(-OR_PI * OR_PI * m2 / (beta * beta)) / m2;
        if (g < tol * 1e-3 && exp(-OR_PI * OR_PI * m2 / (beta * beta)) < tol) continue;
        double Sr = 0, Si = 0;
        for (int i = 0; i < n; i++) {
          const double *base = ph + 2 * (size_t)i * (W0 + W1 + W2);
          const double *px = base + 2 * (kx + nmax[0]), *py = base + 2 * (W0 + ky + nmax[1]),
                       *pz = base + 2 * (W0 + W1 + kz + nmax[2]);
          double xr = px[0] * py[0] - px[1] * py[1], xi = px[0] * py[1] + px[1] * py[0];
          double zr = xr * pz[0] - xi * pz[1], zi = xr * pz[1] + xi * pz[0];
          re_j[2 * i] = zr; re_j[2 * i + 1] = zi;
          Sr += t->charge[i] * zr; Si += t->charge[i] * zi;
        }
        double Em = 2.0 * A * g * (Sr * Sr + Si * Si); /* m and -m */
        et[0] += Em;
        double fac = 2.0 * (1.0 + OR_PI * OR_PI * m2 / (beta * beta)) / m2;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) et[1 + 3 * a + b] += Em * ((a == b) - fac * m[a] * m[b]);
        double pre = 2.0 * 4.0 * OR_PI * A * g;
        for (int i = 0; i < n; i++) {
          /* Im(conj(S) e^{i th_i}) = Sr*si - Si*cr */
          double im = Sr * re_j[2 * i + 1] - Si * re_j[2 * i];
          double c = pre * t->charge[i] * im;
          for (int d = 0; d < 3; d++) ft[3 * i + d] += c * m[d];
        }
      }
    free(re_j);
  }
  for (int th = 0; th < nth; th++) {
    for (int64_t k = 0; k < 3 * (int64_t)n; k++) f[k] += fth[(size_t)th * 3 * n + k];
    E += eth[10 * th];
    for (int k = 0; k < 9; k++) Wv[k] += eth[10 * th + 1 + k];
  }
  e[OR_E_PME] += E;
  for (int k = 0; k < 9; k++) vir[k] += Wv[k];
  free(fth); free(eth); free(ph);
}

void or_forces(or_sys *s, const uint64_t *u, double *f, double *e, double *vir, int flags) {
  int n = s->t.n_atoms;
  double ee[OR_NE] = {0}, vv[9] = {0};
  double *ff = f ? f : (double *)malloc(sizeof(double) * 3 * (size_t)n);
  memset(ff, 0, sizeof(double) * 3 * (size_t)n);
  if (flags & OR_REAL) real_space(s, u, ff, ee, vv);
  if (flags & OR_EXCL) excl_terms(s, u, ff, ee, vv);
  if (flags & OR_BONDED) bonded(s, u, ff, ee, vv);
  if (flags & OR_EWALD) ewald_direct(s, u, ff, ee, vv);
  else if (flags & OR_PME) pme(s, u, ff, ee, vv);
  if ((flags & OR_RESTRAINT) && s->com_on) com_restraint(s, u, ff);
  if (flags & OR_SELF) {
    double q2 = 0;
    for (int i = 0; i < n; i++) q2 += s->t.charge[i] * s->t.charge[i];
    ee[OR_E_EWSELF] = -OR_CC * s->p.beta / sqrt(OR_PI) * q2; /* src/prmread.cu:1734 */
  }
  if (e) memcpy(e, ee, sizeof(ee));
  if (vir) memcpy(vir, vv, sizeof(vv));
  if (!f) free(ff);
}

/* ------------------------------------------------------------------------ */
/* Column/cluster binning and cluster-pair list (the GPU path's definition) */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t z; int32_t id; } zkey;
static int cmp_zkey(const void *a, const void *b) {
  const zkey *x = (const zkey *)a, *y = (const zkey *)b;
  if (x->z != y->z) return x->z < y->z ? -1 : 1;
  return (x->id > y->id) - (x->id < y->id);
}

static inline uint64_t axis_gap(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi) {
  if (alo <= bhi && blo <= ahi) return 0;
  uint64_t d1 = blo - ahi, d2 = alo - bhi;
  return d1 < d2 ? d1 : d2;
}

int64_t or_cluster_pairs(or_sys *s, const uint64_t *u, int32_t nx, int32_t ny,
                         int32_t *atom_column, int32_t *slot_atom, int64_t n_slots_cap,
                         int64_t *n_slots, int32_t *pairs, uint64_t *masks, int64_t cap) {
  const int M = 8;
  int n = s->t.n_atoms;
  int ncol = nx * ny;
  int *cnt = (int *)calloc((size_t)ncol, sizeof(int));
  for (int i = 0; i < n; i++) {
    int cx = (int)(((u128)u[3 * i] * (u128)(uint64_t)nx) >> 64);
    int cy = (int)(((u128)u[3 * i + 1] * (u128)(uint64_t)ny) >> 64);
    atom_column[i] = cy * nx + cx;
    cnt[cy * nx + cx]++;
  }
  int64_t *clu0 = (int64_t *)malloc(sizeof(int64_t) * ((size_t)ncol + 1));
  clu0[0] = 0;
  for (int c = 0; c < ncol; c++) clu0[c + 1] = clu0[c] + (cnt[c] + M - 1) / M;
  int64_t ncl = clu0[ncol];
  *n_slots = ncl * M;
  if (ncl * M > n_slots_cap) { free(cnt); free(clu0); return -1; }
  /* per-column sort by (u_z, id) */
  zkey **col = (zkey **)calloc((size_t)ncol, sizeof(zkey *));
  int *fill = (int *)calloc((size_t)ncol, sizeof(int));
  for (int c = 0; c < ncol; c++) col[c] = (zkey *)malloc(sizeof(zkey) * (size_t)(cnt[c] + 1));
  for (int i = 0; i < n; i++) { int c = atom_column[i]; col[c][fill[c]].z = u[3 * i + 2]; col[c][fill[c]].id = i; fill[c]++; }
  for (int64_t sl = 0; sl < ncl * M; sl++) slot_atom[sl] = -1;
  for (int c = 0; c < ncol; c++) {
    qsort(col[c], (size_t)cnt[c], sizeof(zkey), cmp_zkey);
    for (int k = 0; k < cnt[c]; k++) slot_atom[clu0[c] * M + k] = col[c][k].id;
  }
  /* bounding boxes */
  uint64_t *bb = (uint64_t *)malloc(sizeof(uint64_t) * 6 * (size_t)(ncl + 1));
  for (int64_t I = 0; I < ncl; I++) {
    uint64_t lo[3] = {UINT64_MAX, UINT64_MAX, UINT64_MAX}, hi[3] = {0, 0, 0};
    for (int k = 0; k < M; k++) {
      int a = slot_atom[I * M + k];
      if (a < 0) continue;
      for (int d = 0; d < 3; d++) {
        if (u[3 * a + d] < lo[d]) lo[d] = u[3 * a + d];
        if (u[3 * a + d] > hi[d]) hi[d] = u[3 * a + d];
      }
    }
    for (int d = 0; d < 3; d++) { bb[6 * I + 2 * d] = lo[d]; bb[6 * I + 2 * d + 1] = hi[d]; }
  }
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  double rl2 = s->p.pairlist * s->p.pairlist;
  /* Entry (I, J >= I) exists iff the bounding-box gap is below the pair-list
   * distance and at least one lane pair qualifies: both slots hold atoms, not
   * excluded, i < j inside a self pair, and the FP32 list distance (32-bit top
   * words of the positions, each product and sum rounded to float) is below
   * rl^2 (1 + 1e-6) rounded to float.  The lane mask holds exactly the
   * qualifying pairs.  Per I: (J, mask) in ascending J. */
  const float sc32[3] = {(float)(s->p.box[0] / 4294967296.0), (float)(s->p.box[1] / 4294967296.0),
                         (float)(s->p.box[2] / 4294967296.0)};
  const float rl2f = (float)(s->p.pairlist * s->p.pairlist * (1.0 + 1e-6));
  int64_t *cntI = (int64_t *)calloc((size_t)ncl + 1, sizeof(int64_t));
  int32_t **lj = (int32_t **)calloc((size_t)ncl + 1, sizeof(int32_t *));
  uint64_t **lm = (uint64_t **)calloc((size_t)ncl + 1, sizeof(uint64_t *));
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t I = 0; I < ncl; I++) {
    int64_t c = 0, capI = 64;
    int32_t *vj = (int32_t *)malloc(sizeof(int32_t) * capI);
    uint64_t *vm = (uint64_t *)malloc(sizeof(uint64_t) * capI);
    for (int64_t J = I; J < ncl; J++) {
      double g[3];
      for (int d = 0; d < 3; d++)
        g[d] = (double)axis_gap(bb[6 * I + 2 * d], bb[6 * I + 2 * d + 1], bb[6 * J + 2 * d], bb[6 * J + 2 * d + 1]) * sc[d];
      double d2 = g[0] * g[0] + g[1] * g[1];
      d2 = d2 + g[2] * g[2];
      if (!(d2 < rl2)) continue;
      uint64_t mask = 0;
      for (int l = 0; l < 64; l++) {
        int il = l & 7, jl = l >> 3;
        int a = slot_atom[I * M + il], b = slot_atom[J * M + jl];
        if (a < 0 || b < 0) continue;
        if (I == J && il >= jl) continue;
        if (excluded(s, a, b)) continue;
        float dx = (float)(int32_t)((uint32_t)(u[3 * b] >> 32) - (uint32_t)(u[3 * a] >> 32)) * sc32[0];
        float dy = (float)(int32_t)((uint32_t)(u[3 * b + 1] >> 32) - (uint32_t)(u[3 * a + 1] >> 32)) * sc32[1];
        float dz = (float)(int32_t)((uint32_t)(u[3 * b + 2] >> 32) - (uint32_t)(u[3 * a + 2] >> 32)) * sc32[2];
        float r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
        if (!(r2 < rl2f)) continue;
        mask |= 1ull << l;
      }
      if (!mask) continue;
      if (c == capI) {
        capI *= 2;
        vj = (int32_t *)realloc(vj, sizeof(int32_t) * capI);
        vm = (uint64_t *)realloc(vm, sizeof(uint64_t) * capI);
      }
      vj[c] = (int32_t)J; vm[c] = mask; c++;
    }
    cntI[I + 1] = c; lj[I] = vj; lm[I] = vm;
  }
  for (int64_t I = 0; I < ncl; I++) cntI[I + 1] += cntI[I];
  int64_t np = cntI[ncl];
  if (np <= cap) {
    for (int64_t I = 0; I < ncl; I++)
      for (int64_t k = 0; k < cntI[I + 1] - cntI[I]; k++) {
        int64_t w = cntI[I] + k;
        pairs[2 * w] = (int32_t)I; pairs[2 * w + 1] = lj[I][k]; masks[w] = lm[I][k];
      }
  }
  for (int64_t I = 0; I < ncl; I++) { free(lj[I]); free(lm[I]); }
  free(lj); free(lm);
  for (int c = 0; c < ncol; c++) free(col[c]);
  free(col); free(fill); free(cnt); free(clu0); free(bb); free(cntI);
  return np <= cap ? np : -np;
}

/* ------------------------------------------------------------------------ */
/* Integrator, SHAKE/RATTLE, Maxwell velocities                             */
/* ------------------------------------------------------------------------ */

static int solve_lin(int n, double A[12][12], double *b) {
  /* Gaussian elimination with partial pivoting; b <- A^{-1} b */
  for (int k = 0; k < n; k++) {
    int piv = k;
    for (int r = k + 1; r < n; r++) if (fabs(A[r][k]) > fabs(A[piv][k])) piv = r;
    if (A[piv][k] == 0.0) return -1;
    if (piv != k) {
      for (int c = 0; c < n; c++) { double tmp = A[k][c]; A[k][c] = A[piv][c]; A[piv][c] = tmp; }
      double tmp = b[k]; b[k] = b[piv]; b[piv] = tmp;
    }
    for (int r = k + 1; r < n; r++) {
      double m = A[r][k] / A[k][k];
      for (int c = k; c < n; c++) A[r][c] -= m * A[k][c];
      b[r] -= m * b[k];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    double x = b[k];
    for (int c = k + 1; c < n; c++) x -= A[k][c] * b[c];
    b[k] = x / A[k][k];
  }
  return 0;
}

/* SHAKE of one cluster (Newton iteration on the multipliers, src/shake.cu:20-326),
 * converged when max |r^2 - d^2| / (2 d^2) < tol.  Returns the position
 * corrections dr[natom][3]. */
/* SHAKE of one cluster: new positions unew (box scale sc) against the
 * geometry of uold (box scale sco; the same box except under the barostat) */
static int shake_cluster_ref(const or_sys *s, const or_cluster *cl, const uint64_t *uold, const double sco[3],
                             const uint64_t *unew, const double sc[3], double dr[8][3]) {
  int na = cl->natom, nc = cl->ncons;
  double rn[8][3], ro[8][3], minv[8];
  int a0 = cl->atom[0];
  for (int k = 0; k < na; k++) {
    fx_disp(unew, a0, cl->atom[k], sc, rn[k]);
    fx_disp(uold, a0, cl->atom[k], sco, ro[k]);
    minv[k] = dyn_minv(s, cl->atom[k]);
    dr[k][0] = dr[k][1] = dr[k][2] = 0.0;
  }
  double g[12][3], lam[12] = {0};
  for (int c = 0; c < nc; c++)
    for (int d = 0; d < 3; d++) g[c][d] = ro[cl->c1[c]][d] - ro[cl->c2[c]][d];
  for (int it = 0; it < 1000; it++) {
    double rc_[12][3], sig[12], worst = 0;
    for (int c = 0; c < nc; c++) {
      int i = cl->c1[c], j = cl->c2[c];
      for (int d = 0; d < 3; d++) rc_[c][d] = (rn[i][d] + dr[i][d]) - (rn[j][d] + dr[j][d]);
      sig[c] = dot3(rc_[c], rc_[c]) - cl->d2[c];
      double rel = fabs(sig[c]) / (2.0 * cl->d2[c]);
      if (rel > worst) worst = rel;
    }
    if (worst < s->p.shake_tol) return it;
    double J[12][12];
    for (int c = 0; c < nc; c++)
      for (int e = 0; e < nc; e++) {
        int i = cl->c1[c], j = cl->c2[c], ei = cl->c1[e], ej = cl->c2[e];
        double wi = ((i == ei) - (i == ej)) * minv[i], wj = ((j == ei) - (j == ej)) * minv[j];
        J[c][e] = 2.0 * dot3(rc_[c], g[e]) * (wi - wj);
      }
    for (int c = 0; c < nc; c++) sig[c] = -sig[c];
    if (solve_lin(nc, J, sig)) return -1;
    for (int c = 0; c < nc; c++) lam[c] += sig[c];
    for (int k = 0; k < na; k++) dr[k][0] = dr[k][1] = dr[k][2] = 0.0;
    for (int c = 0; c < nc; c++) {
      int i = cl->c1[c], j = cl->c2[c];
      for (int d = 0; d < 3; d++) { dr[i][d] += lam[c] * g[c][d] * minv[i]; dr[j][d] -= lam[c] * g[c][d] * minv[j]; }
    }
  }
  return -1;
}

static int shake_cluster(const or_sys *s, const or_cluster *cl, const uint64_t *uold, const uint64_t *unew,
                         const double sc[3], double dr[8][3]) {
  return shake_cluster_ref(s, cl, uold, sc, unew, sc, dr);
}

/* RATTLE velocity projection of one cluster (src/shake.cu:330-615): exact solve
 * of r_c . (v_c + dv_c) = 0 for all constraints of the cluster.  With wc, adds
 * the constraint virial of the projection: the velocity change of atom a is
 * minv_a sum_c b_c B_ca r_c, i.e. the half-kick dt/(2 m_a) G_a of the
 * constraint force G_a = (2/dt) sum_c b_c B_ca r_c, whose virial
 * sum_a r_a (x) G_a is (2/dt) sum_c b_c r_c (x) r_c (the constraint-force part
 * of the reference's PCONSTANT virial, src/shake.cu:300-306, there halved and
 * diagonal only). */
static void rattle_cluster(const or_sys *s, const or_cluster *cl, const uint64_t *u, const double sc[3], double *v,
                           double dt, double *wc) {
  int na = cl->natom, nc = cl->ncons;
  double r[8][3], minv[8];
  for (int k = 0; k < na; k++) { fx_disp(u, cl->atom[0], cl->atom[k], sc, r[k]); minv[k] = dyn_minv(s, cl->atom[k]); }
  double rc_[12][3], A[12][12], b[12];
  for (int c = 0; c < nc; c++)
    for (int d = 0; d < 3; d++) rc_[c][d] = r[cl->c1[c]][d] - r[cl->c2[c]][d];
  for (int c = 0; c < nc; c++) {
    int i = cl->c1[c], j = cl->c2[c];
    double vij[3];
    for (int d = 0; d < 3; d++) vij[d] = v[3 * cl->atom[i] + d] - v[3 * cl->atom[j] + d];
    b[c] = -dot3(rc_[c], vij);
    for (int e = 0; e < nc; e++) {
      int ei = cl->c1[e], ej = cl->c2[e];
      double wi = ((i == ei) - (i == ej)) * minv[i], wj = ((j == ei) - (j == ej)) * minv[j];
      A[c][e] = dot3(rc_[c], rc_[e]) * (wi - wj);
    }
  }
  if (solve_lin(nc, A, b)) return;
  for (int c = 0; c < nc; c++) {
    int i = cl->c1[c], j = cl->c2[c];
    for (int d = 0; d < 3; d++) {
      v[3 * cl->atom[i] + d] += b[c] * rc_[c][d] * minv[i];
      v[3 * cl->atom[j] + d] -= b[c] * rc_[c][d] * minv[j];
    }
    if (wc)
      for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) wc[3 * a + d] += 2.0 / dt * b[c] * rc_[c][a] * rc_[c][d];
  }
}

double or_constraint_residual(or_sys *s, const uint64_t *u, double *vres, const double *v) {
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  double worst = 0, vw = 0;
  for (int c = 0; c < s->t.n_constraints; c++) {
    int a = s->t.constraint_atoms[2 * c], b = s->t.constraint_atoms[2 * c + 1];
    double d[3]; fx_disp(u, a, b, sc, d);
    double d2 = s->t.constraint_length[c] * s->t.constraint_length[c];
    double rel = fabs(dot3(d, d) - d2) / (2.0 * d2);
    if (rel > worst) worst = rel;
    if (v) {
      double vv[3] = {v[3 * b] - v[3 * a], v[3 * b + 1] - v[3 * a + 1], v[3 * b + 2] - v[3 * a + 2]};
      double x = fabs(dot3(vv, d)) / sqrt(dot3(d, d));
      if (x > vw) vw = x;
    }
  }
  if (vres) *vres = vw;
  return worst;
}

double or_kinetic(or_sys *s, const double *v) {
  double ke = 0;
  for (int i = 0; i < s->t.n_atoms; i++) ke += s->t.mass[i] * dot3(v + 3 * i, v + 3 * i);
  return 0.5 * ke;
}

double or_temperature(or_sys *s, const double *v) {
  /* T = 2 KE / (RC (3 N_mobile - 6 - Ncons)) (src/evalprops.cu:221-225) */
  return 2.0 * or_kinetic(s, v) / (OR_RC * (3.0 * (s->t.n_atoms - s->n_fixed) - 6.0 - s->n_cons_total));
}

/* Velocity Verlet with SHAKE after the drift and RATTLE after the second kick
 * (SingleStep, src/dynamics.cu:623-820; HalfKickGPU src/mdcuda.cu:27,
 * UpdateCoords :354). */
int or_step(or_sys *s, uint64_t *u, double *v, int nsteps) {
  return or_step_ex(s, u, v, nsteps, NULL, NULL);
}

int or_step_ex(or_sys *s, uint64_t *u, double *v, int nsteps, double *e_out, double *vir_out) {
  int n = s->t.n_atoms;
  double dt = OR_TU * s->p.dt_ps;
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  uint64_t *uold = (uint64_t *)malloc(sizeof(uint64_t) * 3 * (size_t)n);
  double e[OR_NE], vir[9];
  or_forces(s, u, f, e, vir, OR_ALL);
  int rc = 0;
  for (int st = 0; st < nsteps; st++) {
    for (int i = 0; i < n; i++) {
      double hm = 0.5 * dt * dyn_minv(s, i);
      for (int d = 0; d < 3; d++) v[3 * i + d] += hm * f[3 * i + d];
    }
    memcpy(uold, u, sizeof(uint64_t) * 3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) u[3 * i + d] += (uint64_t)(int64_t)llrint(dt * v[3 * i + d] / sc[d]);
    for (int c = 0; c < s->n_clusters; c++) {
      const or_cluster *cl = &s->clusters[c];
      double dr[8][3];
      if (shake_cluster(s, cl, uold, u, sc, dr) < 0) rc = -1;
      for (int k = 0; k < cl->natom; k++)
        for (int d = 0; d < 3; d++) {
          int a = cl->atom[k];
          u[3 * a + d] += (uint64_t)(int64_t)llrint(dr[k][d] / sc[d]);
          v[3 * a + d] += dr[k][d] / dt;
        }
    }
    or_forces(s, u, f, e, vir, OR_ALL);
    for (int i = 0; i < n; i++) {
      double hm = 0.5 * dt * dyn_minv(s, i);
      for (int d = 0; d < 3; d++) v[3 * i + d] += hm * f[3 * i + d];
    }
    for (int c = 0; c < s->n_clusters; c++) rattle_cluster(s, &s->clusters[c], u, sc, v, dt, vir);
  }
  if (e_out) memcpy(e_out, e, sizeof(e));
  if (vir_out) memcpy(vir_out, vir, sizeof(vir));
  free(f); free(uold);
  return rc;
}

/* Andersen barostat, isotropic (NPT with the thermostat of the caller's
 * velocity reassignment).  The piston coordinate is the volume V, with
 * V'' = (P - p_ref) / M: the same piston the reference's reduce_virial forms
 * (src/cell_based_kernels.cu:403-600: (2 KE + W) / (3 V) - pRef PU over the
 * piston mass), whose NPT step is left unfinished there
 * (src/npt_kernels.cu:26-33).  The box is L = l L0, l = cbrt(V / V0).
 *
 * Scaling is molecular for the constraint clusters (Ryckaert & Ciccotti 1983,
 * Andersen's piston with constrained molecules): the box moves each cluster's
 * centre of mass and each atom of no cluster, while a cluster's internal
 * geometry stays as it is, so a box change never strains a constraint.  The
 * piston is driven by the molecular pressure (2 KE_mol + W_mol) / 3V, where
 * KE_mol counts the centre-of-mass motion of clusters and the motion of free
 * atoms, and W_mol = tr sum_i r_i F_i - sum_clusters sum_{i in c}
 * (r_i - R_c) . F_i over the non-constraint forces (constraint forces are
 * internal to their cluster and drop out).  Velocities are peculiar velocities;
 * a box change from l to l' scales a cluster's centre-of-mass velocity (and a
 * free atom's velocity) by l / l'.  The extended energy
 * KE + U + p_ref V + Pi^2 / (2M) is conserved.  One step:
 *   A  Pi += dt/2 (W_mol / 3V - p_ref); V1 = V + dt/2 Pi / M
 *   B  v += dt/2 F / m; box l -> l1 (centre-of-mass velocities and internal
 *      offsets as above); u += dt v at box l1; Pi += dt 2 KE_mol / 3V1
 *   C  V2 = V1 + dt/2 Pi / M; box l1 -> l2
 *   D  SHAKE at box l2 against the cluster geometry before the drift
 *   E  forces and virial at (u, l2)
 *   F  Pi += dt/2 (W_mol / 3V2 - p_ref); v += dt/2 F / m; RATTLE
 * Without clusters every part is an exact flow of the extended Hamiltonian, so
 * the step is symplectic and time reversible. */
static void set_box_scale(or_sys *s, const or_piston *pc, double l) {
  for (int d = 0; d < 3; d++) s->p.box[d] = l * pc->L0[d];
}

/* fixed-point offset of atom a from the cluster's first atom, per axis */
static inline double cl_off(const uint64_t *u, int a0, int a, int d) {
  return (double)(int64_t)(u[3 * a + d] - u[3 * a0 + d]);
}

/* A box change by f = l_old / l_new for every cluster: internal fixed-point
 * offsets from the centre of mass scale by f (their length in A stays), the
 * centre-of-mass velocity scales by f; free atoms' velocities scale by f
 * (their scaled positions stay).  With fkin, accumulates the molecular kinetic
 * energy after the change. */
static void npt_box_change(const or_sys *s, uint64_t *u, double *v, double f, double *ke_mol) {
  const int n = s->t.n_atoms;
  double ke = 0.0;
  for (int c = 0; c < s->n_clusters; c++) {
    const or_cluster *cl = &s->clusters[c];
    const int a0 = cl->atom[0];
    double M = 0.0, off[3] = {0, 0, 0}, P[3] = {0, 0, 0};
    for (int k = 0; k < cl->natom; k++) {
      const int a = cl->atom[k];
      const double m = s->t.mass[a];
      M += m;
      for (int d = 0; d < 3; d++) { off[d] += m * cl_off(u, a0, a, d); P[d] += m * v[3 * a + d]; }
    }
    double Vc[3], dl[8][3];
    for (int d = 0; d < 3; d++) { off[d] /= M; Vc[d] = P[d] / M; }
    for (int k = 0; k < cl->natom; k++)
      for (int d = 0; d < 3; d++) dl[k][d] = cl_off(u, a0, cl->atom[k], d) - off[d];
    for (int k = 0; k < cl->natom; k++) {
      const int a = cl->atom[k];
      for (int d = 0; d < 3; d++) {
        u[3 * a + d] += (uint64_t)(int64_t)llrint((f - 1.0) * dl[k][d]);
        v[3 * a + d] += (f - 1.0) * Vc[d];
      }
    }
    for (int d = 0; d < 3; d++) ke += 0.5 * M * (f * Vc[d]) * (f * Vc[d]);
  }
  for (int i = 0; i < n; i++) {
    if (s->in_cluster[i]) continue;
    for (int d = 0; d < 3; d++) {
      v[3 * i + d] *= f;
      ke += 0.5 * s->t.mass[i] * v[3 * i + d] * v[3 * i + d];
    }
  }
  if (ke_mol) *ke_mol = ke;
}

/* sum over clusters of sum_{i in c} (r_i - R_c) . F_i (forces f at positions u, box as set) */
static double npt_intra_virial(const or_sys *s, const uint64_t *u, const double *f) {
  const double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  double w = 0.0;
  for (int c = 0; c < s->n_clusters; c++) {
    const or_cluster *cl = &s->clusters[c];
    const int a0 = cl->atom[0];
    double M = 0.0, R[3] = {0, 0, 0};
    double r[8][3];
    for (int k = 0; k < cl->natom; k++) {
      const int a = cl->atom[k];
      const double m = s->t.mass[a];
      fx_disp(u, a0, a, sc, r[k]);
      M += m;
      for (int d = 0; d < 3; d++) R[d] += m * r[k][d];
    }
    for (int d = 0; d < 3; d++) R[d] /= M;
    for (int k = 0; k < cl->natom; k++)
      for (int d = 0; d < 3; d++) w += (r[k][d] - R[d]) * f[3 * cl->atom[k] + d];
  }
  return w;
}

int or_step_npt(or_sys *s, or_piston *pc, uint64_t *u, double *v, int nsteps, double *e_out, double *vir_out) {
  int n = s->t.n_atoms;
  double dt = OR_TU * s->p.dt_ps;
  const double V0 = pc->L0[0] * pc->L0[1] * pc->L0[2], M = pc->mass;
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  uint64_t *uold = (uint64_t *)malloc(sizeof(uint64_t) * 3 * (size_t)n);
  double e[OR_NE], vir[9];
  double V = pc->V, l = cbrt(V / V0);
  set_box_scale(s, pc, l);
  or_forces(s, u, f, e, vir, OR_ALL);
  int rc = 0;
  for (int st = 0; st < nsteps; st++) {
    /* A */
    pc->Pi += 0.5 * dt * (pc->Wtr / (3.0 * V) - pc->p_ref);
    const double V1 = V + 0.5 * dt * pc->Pi / M, l1 = cbrt(V1 / V0);
    /* B: kick, box l -> l1, drift at l1, piston kick with the molecular KE */
    for (int i = 0; i < n; i++) {
      double hm = 0.5 * dt * dyn_minv(s, i);
      for (int d = 0; d < 3; d++) v[3 * i + d] += hm * f[3 * i + d];
    }
    const double sc0[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
    memcpy(uold, u, sizeof(uint64_t) * 3 * (size_t)n);
    double ke1 = 0.0;
    npt_box_change(s, u, v, l / l1, &ke1);
    set_box_scale(s, pc, l1);
    double sc1[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) u[3 * i + d] += (uint64_t)(int64_t)llrint(dt * v[3 * i + d] / sc1[d]);
    pc->Pi += dt * 2.0 * ke1 / (3.0 * V1);
    /* C */
    const double V2 = V1 + 0.5 * dt * pc->Pi / M, l2 = cbrt(V2 / V0);
    npt_box_change(s, u, v, l1 / l2, NULL);
    set_box_scale(s, pc, l2);
    double sc2[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
    /* D: the cluster geometry before the drift is the old positions at the old box */
    for (int c = 0; c < s->n_clusters; c++) {
      const or_cluster *cl = &s->clusters[c];
      double dr[8][3];
      if (shake_cluster_ref(s, cl, uold, sc0, u, sc2, dr) < 0) rc = -1;
      for (int k = 0; k < cl->natom; k++)
        for (int d = 0; d < 3; d++) {
          int a = cl->atom[k];
          u[3 * a + d] += (uint64_t)(int64_t)llrint(dr[k][d] / sc2[d]);
          v[3 * a + d] += dr[k][d] / dt;
        }
    }
    /* E */
    or_forces(s, u, f, e, vir, OR_ALL);
    /* F */
    pc->Wtr = vir[0] + vir[4] + vir[8] - npt_intra_virial(s, u, f);
    pc->Pi += 0.5 * dt * (pc->Wtr / (3.0 * V2) - pc->p_ref);
    for (int i = 0; i < n; i++) {
      double hm = 0.5 * dt * dyn_minv(s, i);
      for (int d = 0; d < 3; d++) v[3 * i + d] += hm * f[3 * i + d];
    }
    for (int c = 0; c < s->n_clusters; c++) rattle_cluster(s, &s->clusters[c], u, sc2, v, dt, vir);
    V = V2;
    l = l2;
  }
  pc->V = V;
  if (e_out) memcpy(e_out, e, sizeof(e));
  if (vir_out) memcpy(vir_out, vir, sizeof(vir));
  free(f); free(uold);
  return rc;
}

/* The piston's molecular virial W_mol at positions u (box as set). */
double or_virial_trace(or_sys *s, const uint64_t *u) {
  int n = s->t.n_atoms;
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double e[OR_NE], vir[9];
  or_forces(s, u, f, e, vir, OR_ALL);
  const double w = vir[0] + vir[4] + vir[8] - npt_intra_virial(s, u, f);
  free(f);
  return w;
}

void or_set_box(or_sys *s, const double box[3]) {
  for (int d = 0; d < 3; d++) s->p.box[d] = box[d];
}

void or_get_box(const or_sys *s, double box[3]) {
  for (int d = 0; d < 3; d++) box[d] = s->p.box[d];
}

/* SHAKE of every constraint cluster of positions u, with uold as the
 * reference (the minimiser's constraint step after each move). */
int or_shake_positions(or_sys *s, const uint64_t *uold, uint64_t *u) {
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  int rc = 0;
  for (int c = 0; c < s->n_clusters; c++) {
    const or_cluster *cl = &s->clusters[c];
    double dr[8][3];
    if (shake_cluster(s, cl, uold, u, sc, dr) < 0) rc = -1;
    for (int k = 0; k < cl->natom; k++)
      for (int d = 0; d < 3; d++) u[3 * cl->atom[k] + d] += (uint64_t)(int64_t)llrint(dr[k][d] / sc[d]);
  }
  return rc;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void or_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  uint32_t k[2] = {key_in[0], key_in[1]};
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u;
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Maxwell velocities: v_d = sqrt(RC T / m) * N(0,1) (src/init.cu:80-118),
 * counter-based so that each atom's draw is independent of thread layout;
 * then total momentum removed as in src/init.cu:94-118 and RATTLE projection. */
void or_maxwell(or_sys *s, const uint64_t *u, double T, uint64_t seed, uint64_t counter, double *v) {
  int n = s->t.n_atoms;
  double P[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    uint32_t ctr[4] = {(uint32_t)i, 0u, (uint32_t)counter, (uint32_t)(counter >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, r[4];
    or_philox(ctr, key, r);
    double u1 = ((double)r[0] + 0.5) * (1.0 / 4294967296.0), u2 = ((double)r[1] + 0.5) * (1.0 / 4294967296.0);
    double u3 = ((double)r[2] + 0.5) * (1.0 / 4294967296.0), u4 = ((double)r[3] + 0.5) * (1.0 / 4294967296.0);
    double m = s->t.mass[i], vm = (s->mobile && !s->mobile[i]) ? 0.0 : sqrt(OR_RC * T / m);
    double ra = sqrt(-2.0 * log(u1)), rb = sqrt(-2.0 * log(u3));
    v[3 * i] = vm * ra * cos(2.0 * OR_PI * u2);
    v[3 * i + 1] = vm * ra * sin(2.0 * OR_PI * u2);
    v[3 * i + 2] = vm * rb * cos(2.0 * OR_PI * u4);
    for (int d = 0; d < 3; d++) P[d] += m * v[3 * i + d];
  }
  /* momentum removed over the mobile atoms */
  int nm = n - s->n_fixed;
  for (int i = 0; i < n; i++)
    if (!s->mobile || s->mobile[i])
      for (int d = 0; d < 3; d++) v[3 * i + d] -= P[d] / nm / s->t.mass[i];
  double sc[3] = {s->p.box[0] * TWO_POW_M64, s->p.box[1] * TWO_POW_M64, s->p.box[2] * TWO_POW_M64};
  for (int c = 0; c < s->n_clusters; c++) rattle_cluster(s, &s->clusters[c], u, sc, v, 1.0, NULL);
}
