// fz_engine.cpp -- host runtime of libfenzi.
//
// Replaces the reference's InitHostMem / InitMem / InitDeviceConstants /
// GenExcl_minimal_new / build_constraints (src/prmread.cu:86,1303,2928;
// src/mem.cu:49; src/cell_based_kernels.cu:1831) and the step orchestration of
// SingleStep / ComputeAccelGPU (src/dynamics.cu:623, src/mdcuda.cu:557).
// Capacities that the reference fixes at compile time (CELL_ATOMS, MAXNB,
// MAXLATTICE; src/configure:89-137) are sized from the input at run time and
// grown on overflow, after which the interrupted steps are replayed from a
// device-side snapshot.
#include "fz_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace fz {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
const char* last_error() { return g_err.c_str(); }

static const double kPi = 3.14159265358979323846;
static const double kCC = 332.0716, kTU = 20.4548, kRC = 0.001987;
static const double kPU = 1.4586e-5;  // atm -> kcal/mol/A^3 (PU, src/defs.in.h:105)
static const double kTwo64 = 18446744073709551616.0;

void launch_zero_forces(const Launch& L);

// ---------------------------------------------------------------------------
// topology processing

static bool smooth235(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

static int uf_find(std::vector<int>& p, int x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

static inline uint64_t pkey(int a, int b, int n) {
  int lo = std::min(a, b), hi = std::max(a, b);
  return (uint64_t)lo * (uint64_t)n + (uint64_t)hi;
}

// Constraint clusters (connected components of the constraint graph, at most
// 4 atoms and 3 constraints each; src/prmread.cu:1192-1303).  Constraints whose
// atoms are all fixed are left out: they cannot move and would make the
// cluster's SHAKE matrix singular.
static int build_ccl(HostTopo& h) {
  const int n = h.n;
  h.ccl_atoms.clear();
  h.ccl_cons.clear();
  h.ccl_d2.clear();
  h.n_cons_active = 0;
  std::vector<int> act;
  for (int c = 0; c < (int)h.cons_ab.size(); c++)
    if (!(h.fixed[h.cons_ab[c].x] && h.fixed[h.cons_ab[c].y])) act.push_back(c);
  h.n_cons_active = (int)act.size();
  std::vector<uint8_t> in_cl(n, 0);
  for (int c : act) in_cl[h.cons_ab[c].x] = in_cl[h.cons_ab[c].y] = 1;
  h.free_atoms.clear();
  for (int i = 0; i < n; i++)
    if (!in_cl[i]) h.free_atoms.push_back(i);
  if (act.empty()) return FZ_OK;
  std::vector<int> par(n);
  std::iota(par.begin(), par.end(), 0);
  for (int c : act) {
    int ra = uf_find(par, h.cons_ab[c].x), rb = uf_find(par, h.cons_ab[c].y);
    if (ra != rb) par[ra] = rb;
  }
  std::vector<int> cid(n, -1);
  struct Cl { int at[4]; int na; int c1[3], c2[3]; double d2[3]; int nc; };
  std::vector<Cl> cls;
  for (int c : act) {
    const int a = h.cons_ab[c].x, b = h.cons_ab[c].y;
    int r = uf_find(par, a);
    if (cid[r] < 0) {
      cid[r] = (int)cls.size();
      Cl z;
      memset(&z, 0, sizeof z);
      for (int k = 0; k < 4; k++) z.at[k] = -1;
      cls.push_back(z);
    }
    Cl& cl = cls[cid[r]];
    int la = -1, lb = -1;
    for (int k = 0; k < cl.na; k++) { if (cl.at[k] == a) la = k; if (cl.at[k] == b) lb = k; }
    if (la < 0) { if (cl.na >= 4) return set_error(FZ_ERR_ARG, "constraint cluster with more than 4 atoms"); la = cl.na; cl.at[cl.na++] = a; }
    if (lb < 0) { if (cl.na >= 4) return set_error(FZ_ERR_ARG, "constraint cluster with more than 4 atoms"); lb = cl.na; cl.at[cl.na++] = b; }
    if (cl.nc >= 3) return set_error(FZ_ERR_ARG, "constraint cluster with more than 3 constraints");
    cl.c1[cl.nc] = la; cl.c2[cl.nc] = lb;
    cl.d2[cl.nc] = h.cons_len[c] * h.cons_len[c];
    cl.nc++;
  }
  for (auto& cl : cls) {
    h.ccl_atoms.push_back(make_int4(cl.at[0], cl.at[1], cl.at[2], cl.at[3]));
    int p[3] = {0, 0, 0};
    for (int k = 0; k < cl.nc; k++) p[k] = cl.c1[k] | (cl.c2[k] << 8);
    h.ccl_cons.push_back(make_int4(p[0], p[1], p[2], cl.nc));
    h.ccl_d2.push_back(make_double4(cl.d2[0], cl.nc > 1 ? cl.d2[1] : 1.0, cl.nc > 2 ? cl.d2[2] : 1.0, 0.0));
  }
  return FZ_OK;
}

static int process_topology(const fz_topology* t, HostTopo& h) {
  const int n = t->n_atoms;
  if (n <= 0) return set_error(FZ_ERR_ARG, "n_atoms must be positive");
  if (t->n_types <= 0 || t->n_types > 256) return set_error(FZ_ERR_ARG, "n_types must be in 1..256");
  h.n = n;
  h.ntypes = t->n_types;
  h.q.assign(t->charge, t->charge + n);
  h.mass.assign(t->mass, t->mass + n);
  h.type.assign(t->type, t->type + n);
  h.minv.resize(n);
  h.q2sum = 0.0;
  for (int i = 0; i < n; i++) {
    if (!(h.mass[i] > 0.0)) return set_error(FZ_ERR_ARG, "atom %d has non-positive mass", i);
    if (h.type[i] < 0 || h.type[i] >= t->n_types) return set_error(FZ_ERR_ARG, "atom %d has type out of range", i);
    h.minv[i] = 1.0 / h.mass[i];
    h.q2sum += h.q[i] * h.q[i];
  }
  const int T2 = t->n_types * t->n_types;
  h.lj_eps.assign(t->lj_eps, t->lj_eps + T2);
  h.lj_sig2.assign(t->lj_sig2, t->lj_sig2 + T2);
  h.lj14_eps.assign(t->lj14_eps, t->lj14_eps + T2);
  h.lj14_sig2.assign(t->lj14_sig2, t->lj14_sig2 + T2);
  auto chk = [&](int a) { return a >= 0 && a < n; };
  for (int i = 0; i < t->n_bonds; i++) {
    int a = t->bond_atoms[2 * i], b = t->bond_atoms[2 * i + 1];
    if (!chk(a) || !chk(b)) return set_error(FZ_ERR_ARG, "bond %d atom out of range", i);
    h.bond_ij.push_back(make_int2(a, b));
    h.bond_p.push_back(make_double2(t->bond_k[i], t->bond_r0[i]));
  }
  for (int i = 0; i < t->n_angles; i++) {
    int a = t->angle_atoms[3 * i], b = t->angle_atoms[3 * i + 1], c = t->angle_atoms[3 * i + 2];
    if (!chk(a) || !chk(b) || !chk(c)) return set_error(FZ_ERR_ARG, "angle %d atom out of range", i);
    h.angle_ijk.push_back(make_int4(a, b, c, 0));
    h.angle_p.push_back(make_double4(t->angle_k[i], t->angle_theta0[i], t->ub_k ? t->ub_k[i] : 0.0,
                                     t->ub_s0 ? t->ub_s0[i] : 0.0));
  }
  for (int i = 0; i < t->n_dihedrals; i++) {
    const int32_t* a = t->dihedral_atoms + 4 * i;
    for (int k = 0; k < 4; k++)
      if (!chk(a[k])) return set_error(FZ_ERR_ARG, "dihedral %d atom out of range", i);
    h.dihe_ijkl.push_back(make_int4(a[0], a[1], a[2], a[3]));
    h.dihe_p.push_back(make_double2(t->dihedral_k[i], t->dihedral_delta[i]));
    h.dihe_n.push_back(t->dihedral_n[i]);
  }
  for (int i = 0; i < t->n_impropers; i++) {
    const int32_t* a = t->improper_atoms + 4 * i;
    for (int k = 0; k < 4; k++)
      if (!chk(a[k])) return set_error(FZ_ERR_ARG, "improper %d atom out of range", i);
    h.impr_ijkl.push_back(make_int4(a[0], a[1], a[2], a[3]));
    h.impr_k.push_back(t->improper_k[i]);
  }
  // exclusions (src/prmread.cu:2928-3062): 1-2 bonds, 1-3 angle ends -> Ewald
  // correction set; 1-4 dihedral ends -> unique 1-4 pairs not already 1-2/1-3
  std::vector<uint64_t> ew, p14;
  for (auto& b : h.bond_ij)
    if (b.x != b.y) ew.push_back(pkey(b.x, b.y, n));
  for (auto& a : h.angle_ijk)
    if (a.x != a.z) ew.push_back(pkey(a.x, a.z, n));
  std::sort(ew.begin(), ew.end());
  ew.erase(std::unique(ew.begin(), ew.end()), ew.end());
  for (auto& d : h.dihe_ijkl) {
    if (d.x == d.w) continue;
    uint64_t k = pkey(d.x, d.w, n);
    if (!std::binary_search(ew.begin(), ew.end(), k)) p14.push_back(k);
  }
  std::sort(p14.begin(), p14.end());
  p14.erase(std::unique(p14.begin(), p14.end()), p14.end());
  for (uint64_t k : ew) h.ew_ij.push_back(make_int2((int)(k / n), (int)(k % n)));
  for (uint64_t k : p14) h.p14_ij.push_back(make_int2((int)(k / n), (int)(k % n)));
  h.excl_bits.assign(n, 0ull);
  std::vector<std::vector<int>> ex(n);
  bool wide = false;
  auto add_ex = [&](int a, int b) {
    ex[a].push_back(b);
    ex[b].push_back(a);
    int off = b - a;
    if (off >= -32 && off <= 31) h.excl_bits[a] |= 1ull << (off + 32); else wide = true;
    off = a - b;
    if (off >= -32 && off <= 31) h.excl_bits[b] |= 1ull << (off + 32); else wide = true;
  };
  for (auto& p : h.ew_ij) add_ex(p.x, p.y);
  for (auto& p : h.p14_ij) add_ex(p.x, p.y);
  h.use_excl_list = wide ? 1 : 0;
  h.excl_off.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    std::sort(ex[i].begin(), ex[i].end());
    h.excl_off[i + 1] = h.excl_off[i] + (int)ex[i].size();
  }
  h.excl_list.clear();
  h.excl_list.reserve(h.excl_off[n] + 1);
  for (int i = 0; i < n; i++) h.excl_list.insert(h.excl_list.end(), ex[i].begin(), ex[i].end());
  if (h.excl_list.empty()) h.excl_list.push_back(-1);
  // constraint clusters (src/prmread.cu:1192-1303)
  h.n_constraints = t->n_constraints;
  h.cons_ab.clear();
  h.cons_len.clear();
  for (int c = 0; c < t->n_constraints; c++) {
    int a = t->constraint_atoms[2 * c], b = t->constraint_atoms[2 * c + 1];
    if (!chk(a) || !chk(b) || a == b) return set_error(FZ_ERR_ARG, "constraint %d atoms invalid", c);
    h.cons_ab.push_back(make_int2(a, b));
    h.cons_len.push_back(t->constraint_length[c]);
  }
  h.fixed.assign(n, 0);
  h.n_fixed = 0;
  FZ_TRY(build_ccl(h));
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// System

System::~System() { free_device(); }

void System::free_device() {
  if (!allocs.empty() || !nb_allocs.empty()) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (auto& a : allocs) (void)hipFree(a.first);
    for (auto& a : nb_allocs) (void)hipFree(a.first);
    (void)hipSetDevice(cur);
  }
  allocs.clear();
  nb_allocs.clear();
  bytes = 0;
}

void System::free_neighbor() {
  for (auto& a : nb_allocs) {
    (void)hipFree(a.first);
    bytes -= a.second;
  }
  nb_allocs.clear();
}

void System::dfree(void* p) {
  if (!p) return;
  for (size_t k = 0; k < allocs.size(); k++)
    if (allocs[k].first == p) {
      (void)hipFree(p);
      bytes -= allocs[k].second;
      allocs.erase(allocs.begin() + (long)k);
      return;
    }
}

template <class T>
int System::dalloc(T** p, size_t count, bool zero) {
  size_t b = std::max<size_t>(1, count) * sizeof(T);
  b = (b + 255) & ~(size_t)255;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, b);
  if (e != hipSuccess) return set_error(FZ_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
  (alloc_nb ? nb_allocs : allocs).emplace_back(q, b);
  bytes += b;
  if (zero) FZ_HIP(hipMemset(q, 0, b));
  *p = (T*)q;
  return FZ_OK;
}

template <class T>
static int up(T* dst, const std::vector<T>& v) {
  if (v.empty()) return FZ_OK;
  FZ_HIP(hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return FZ_OK;
}

int System::init(const fz_topology* t, const fz_params* p, int precision, int dev) {
  device = dev;
  mixed = precision == FZ_MIXED ? 1 : 0;
  if (precision != FZ_FP64 && precision != FZ_MIXED) return set_error(FZ_ERR_ARG, "unknown precision %d", precision);
  par = *p;
  FZ_TRY(process_topology(t, topo));
  for (int d = 0; d < 3; d++) {
    if (!(par.box[d] > 0)) return set_error(FZ_ERR_ARG, "box must be positive");
    if (!(par.cutoff < 0.5 * par.box[d])) return set_error(FZ_ERR_ARG, "cutoff %.3f must be below half the box (%.3f)", par.cutoff, par.box[d]);
    if (!smooth235(par.pme_grid[d]) || par.pme_grid[d] > MAX_FFT || par.pme_grid[d] < 4)
      return set_error(FZ_ERR_ARG, "PME grid %d must be 2,3,5-smooth and in 4..%d", par.pme_grid[d], MAX_FFT);
  }
  if (par.pme_order != 0 && par.pme_order != 4) return set_error(FZ_ERR_ARG, "only PME order 4 is supported");
  if (!(par.cuton >= 0 && par.cuton < par.cutoff)) return set_error(FZ_ERR_ARG, "cuton must be in [0, cutoff)");
  if (!(par.pairlist >= par.cutoff)) return set_error(FZ_ERR_ARG, "pairlistdist must be >= cutoff");
  if (!(par.beta > 0)) return set_error(FZ_ERR_ARG, "beta (KAPPa) must be positive");
  if (!(par.dt_ps > 0)) par.dt_ps = 0.001;
  if (!(par.shake_tol > 0)) par.shake_tol = 1e-8;
  self_energy = -kCC * par.beta / std::sqrt(kPi) * topo.q2sum;
  return alloc_device();
}

int System::alloc_device() {
  const HostTopo& h = topo;
  SysDev& d = dv;
  memset(&d, 0, sizeof d);
  const int n = h.n;
  d.n_atoms = n;
  d.n_types = h.ntypes;
  d.mixed = mixed;
  {  // odd multiplier near 0.618 n, coprime to n
    unsigned long long m = (unsigned long long)(0.6180339887 * n) | 1ull;
    while (std::gcd(m, (unsigned long long)n) != 1ull) m += 2ull;
    d.spread_mult = m;
  }
  const double V = par.box[0] * par.box[1] * par.box[2];
  for (int k = 0; k < 3; k++) {
    d.L[k] = par.box[k];
    d.sc[k] = par.box[k] / kTwo64;
    d.isc[k] = kTwo64 / par.box[k];
    d.sc32[k] = (float)(par.box[k] / 4294967296.0);
    d.K[k] = par.pme_grid[k];
  }
  d.Kxh = d.K[0] / 2 + 1;
  d.rc2 = par.cutoff * par.cutoff;
  d.ron2 = par.cuton * par.cuton;
  d.rl2 = par.pairlist * par.pairlist;
  d.rl2f = (float)(par.pairlist * par.pairlist * (1.0 + 1e-6));
  d.zwin = (unsigned long long)(par.pairlist / d.sc[2] * (1.0 + 1e-9)) + 64ull;
  double half_skin = 0.5 * (par.pairlist - par.cutoff);
  d.skin2 = half_skin * half_skin;
  d.rl = par.pairlist;
  d.rc = par.cutoff;
  d.beta = par.beta;
  d.sw_inv = 1.0 / std::pow(d.rc2 - d.ron2, 3);
  d.dt = kTU * par.dt_ps;
  d.shake_tol = par.shake_tol;
  d.pme_pre = kCC / (kPi * V);
  d.pi2_beta2 = kPi * kPi / (par.beta * par.beta);
  d.use_excl_list = h.use_excl_list;
  d.n_ccl = (int)h.ccl_atoms.size();
  d.n_free = (int)h.free_atoms.size();
  d.n_bonds = (int)h.bond_ij.size();
  d.n_angles = (int)h.angle_ijk.size();
  d.n_diheds = (int)h.dihe_ijkl.size();
  d.n_imprs = (int)h.impr_ijkl.size();
  d.n_ew = (int)h.ew_ij.size();
  d.n_p14 = (int)h.p14_ij.size();
  // per-atom state
  FZ_TRY(dalloc(&d.ux, n)); FZ_TRY(dalloc(&d.uy, n)); FZ_TRY(dalloc(&d.uz, n));
  FZ_TRY(dalloc(&d.rx, n)); FZ_TRY(dalloc(&d.ry, n)); FZ_TRY(dalloc(&d.rz, n));
  FZ_TRY(dalloc(&d.ox, n)); FZ_TRY(dalloc(&d.oy, n)); FZ_TRY(dalloc(&d.oz, n));
  FZ_TRY(dalloc(&d.vx, n)); FZ_TRY(dalloc(&d.vy, n)); FZ_TRY(dalloc(&d.vz, n));
  FZ_TRY(dalloc(&d.q, n)); FZ_TRY(dalloc(&d.minv, n)); FZ_TRY(dalloc(&d.type, n));
  FZ_TRY(up(d.q, h.q)); FZ_TRY(up(d.minv, h.minv)); FZ_TRY(up(d.type, h.type));
  FZ_TRY(dalloc(&d.mdyn, n));
  FZ_TRY(dalloc(&d.tick, 1));
  FZ_TRY(upload_mdyn());
  FZ_TRY(dalloc(&d.com_grp, n));
  FZ_TRY(dalloc(&d.atom_slot, n)); FZ_TRY(dalloc(&d.atom_col, n));
  FZ_TRY(dalloc(&d.n_clusters, 2)); FZ_TRY(dalloc(&d.n_pairs, 1));
  FZ_TRY(dalloc(&d.excl_bits, n)); FZ_TRY(up(d.excl_bits, h.excl_bits));
  FZ_TRY(dalloc(&d.excl_off, n + 1)); FZ_TRY(up(d.excl_off, h.excl_off));
  FZ_TRY(dalloc(&d.excl_list, h.excl_list.size())); FZ_TRY(up(d.excl_list, h.excl_list));
  FZ_TRY(dalloc(&d.bond_ij, h.bond_ij.size())); FZ_TRY(up(d.bond_ij, h.bond_ij));
  FZ_TRY(dalloc(&d.bond_p, h.bond_p.size())); FZ_TRY(up(d.bond_p, h.bond_p));
  FZ_TRY(dalloc(&d.angle_ijk, h.angle_ijk.size())); FZ_TRY(up(d.angle_ijk, h.angle_ijk));
  FZ_TRY(dalloc(&d.angle_p, h.angle_p.size())); FZ_TRY(up(d.angle_p, h.angle_p));
  FZ_TRY(dalloc(&d.dihe_ijkl, h.dihe_ijkl.size())); FZ_TRY(up(d.dihe_ijkl, h.dihe_ijkl));
  FZ_TRY(dalloc(&d.dihe_p, h.dihe_p.size())); FZ_TRY(up(d.dihe_p, h.dihe_p));
  FZ_TRY(dalloc(&d.dihe_n, h.dihe_n.size())); FZ_TRY(up(d.dihe_n, h.dihe_n));
  FZ_TRY(dalloc(&d.impr_ijkl, h.impr_ijkl.size())); FZ_TRY(up(d.impr_ijkl, h.impr_ijkl));
  FZ_TRY(dalloc(&d.impr_k, h.impr_k.size())); FZ_TRY(up(d.impr_k, h.impr_k));
  FZ_TRY(dalloc(&d.ew_ij, h.ew_ij.size())); FZ_TRY(up(d.ew_ij, h.ew_ij));
  FZ_TRY(dalloc(&d.p14_ij, h.p14_ij.size())); FZ_TRY(up(d.p14_ij, h.p14_ij));
  {
    const int T2 = h.ntypes * h.ntypes;
    std::vector<float2> l32(T2);
    std::vector<double2> l64(T2), l14(T2);
    for (int k = 0; k < T2; k++) {
      l32[k] = make_float2((float)h.lj_eps[k], (float)h.lj_sig2[k]);
      l64[k] = make_double2(h.lj_eps[k], h.lj_sig2[k]);
      l14[k] = make_double2(h.lj14_eps[k], h.lj14_sig2[k]);
    }
    FZ_TRY(dalloc(&d.lj32, T2)); FZ_TRY(up(d.lj32, l32));
    FZ_TRY(dalloc(&d.lj64, T2)); FZ_TRY(up(d.lj64, l64));
    FZ_TRY(dalloc(&d.lj14, T2)); FZ_TRY(up(d.lj14, l14));
    // combination-rule test (pairs with zero well depth carry no sigma information)
    const int T = h.ntypes;
    bool sep = true;
    for (int i = 0; i < T && sep; i++)
      for (int j = 0; j < T && sep; j++) {
        const double eii = h.lj_eps[i * T + i], ejj = h.lj_eps[j * T + j], eij = h.lj_eps[i * T + j];
        if (!(eii >= 0.0 && ejj >= 0.0)) { sep = false; break; }
        const double ec = std::sqrt(eii * ejj);
        if (std::fabs(eij - ec) > 1e-12 * std::max(1.0, std::fabs(ec))) sep = false;
        if (eij == 0.0) continue;
        const double sc = 0.5 * (std::sqrt(h.lj_sig2[i * T + i]) + std::sqrt(h.lj_sig2[j * T + j]));
        if (std::fabs(std::sqrt(h.lj_sig2[i * T + j]) - sc) > 1e-12 * std::max(1.0, sc)) sep = false;
      }
    std::vector<float2> ah(T);
    for (int t = 0; t < T; t++)
      ah[t] = make_float2((float)std::sqrt(std::max(0.0, h.lj_eps[t * T + t])),
                          (float)(0.5 * std::sqrt(std::max(0.0, h.lj_sig2[t * T + t]))));
    d.lj_sep = sep ? 1 : 0;
    FZ_TRY(dalloc(&d.lj_ah, T)); FZ_TRY(up(d.lj_ah, ah));
  }
  // constraints
  FZ_TRY(dalloc(&d.ccl_atoms, h.ccl_atoms.size())); FZ_TRY(up(d.ccl_atoms, h.ccl_atoms));
  FZ_TRY(dalloc(&d.ccl_cons, h.ccl_cons.size())); FZ_TRY(up(d.ccl_cons, h.ccl_cons));
  FZ_TRY(dalloc(&d.ccl_d2, h.ccl_d2.size())); FZ_TRY(up(d.ccl_d2, h.ccl_d2));
  FZ_TRY(dalloc(&d.free_atoms, h.free_atoms.size())); FZ_TRY(up(d.free_atoms, h.free_atoms));
  FZ_TRY(dalloc(&d.br_atom, n, false));
  // PME
  {
    const size_t K3 = (size_t)d.K[0] * d.K[1] * d.K[2];
    const size_t KH = (size_t)d.Kxh * d.K[1] * d.K[2];
    const size_t csz = mixed ? 8 : 16;
    void* p;
    FZ_TRY(dalloc((unsigned long long**)&p, K3)); d.Qacc = p;
    FZ_TRY(dalloc((char**)&p, KH * csz)); d.Qc = p;
    FZ_TRY(dalloc((char**)&p, K3 * (mixed ? 4 : 8))); d.phi = p;
    for (int k = 0; k < 3; k++) d.nbr[k] = (d.K[k] + 7) / 8;
    d.n_bricks = d.nbr[0] * d.nbr[1] * d.nbr[2];
    FZ_TRY(dalloc(&d.br_cnt, d.n_bricks));
    FZ_TRY(dalloc(&d.br_off, d.n_bricks + 1));
    FZ_TRY(dalloc((char**)&p, (size_t)d.n_atoms * 16 * (mixed ? 4 : 8), false)); d.br_rec = p;
    FZ_TRY(dalloc((unsigned long long**)&p, (size_t)d.n_bricks * 1331, false)); d.br_part = p;
    const int KS = d.K[0] + d.K[1] + d.K[2];
    std::vector<double> bsp(KS), ex(KS), mv(KS);
    std::vector<double2> tw64(KS);
    std::vector<float2> tw32(KS);
    int off = 0;
    for (int ax = 0; ax < 3; ax++) {
      const int K = d.K[ax];
      for (int j = 0; j < K; j++) {
        const int jj = j <= K / 2 ? j : j - K;
        const double m = jj / par.box[ax];
        const double c = std::cos(2.0 * kPi * j / K);
        bsp[off + j] = 9.0 / ((2.0 + c) * (2.0 + c));
        ex[off + j] = std::exp(-kPi * kPi * m * m / (par.beta * par.beta));
        mv[off + j] = m;
        const double a = -2.0 * kPi * j / K;
        tw64[off + j] = make_double2(std::cos(a), std::sin(a));
        tw32[off + j] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
      off += K;
    }
    FZ_TRY(dalloc(&d.pme_bsp, KS)); FZ_TRY(up(d.pme_bsp, bsp));
    FZ_TRY(dalloc(&d.pme_ex, KS)); FZ_TRY(up(d.pme_ex, ex));
    FZ_TRY(dalloc(&d.pme_m, KS)); FZ_TRY(up(d.pme_m, mv));
    if (mixed) { float2* t; FZ_TRY(dalloc(&t, KS)); FZ_TRY(up(t, tw32)); d.tw = t; }
    else { double2* t; FZ_TRY(dalloc(&t, KS)); FZ_TRY(up(t, tw64)); d.tw = t; }
    // G(m) for the box at creation (the z pass forms it per point once a barostat moves the box)
    FZ_TRY(dalloc(&d.pme_g, (size_t)d.K[2] * d.K[1] * d.Kxh, false));
    launch_pme_gtable(d, nullptr);
    FZ_HIP(hipDeviceSynchronize());
  }
  FZ_TRY(dalloc(&d.acc, A_NSLOT));
  FZ_TRY(dalloc(&d.acc_ev, (size_t)EV_SPREAD * EV_W));
  FZ_TRY(dalloc((unsigned long long**)&d.Fa, (size_t)3 * n));
  FZ_TRY(dalloc((unsigned long long**)&d.Fg, (size_t)3 * n));
  FZ_TRY(dalloc(&d.flags, F_NSLOT));
  FZ_TRY(dalloc(&d.mstep, 2));
  FZ_TRY(dalloc(&d.mH, (size_t)3 * n));
  FZ_TRY(dalloc(&d.mF, (size_t)3 * n));
  FZ_TRY(dalloc(&snap, (size_t)9 * n + (F_NSLOT * sizeof(int) + 7) / 8));
  d.snap = snap;
  return set_column_grid(par.column_grid[0], par.column_grid[1], 1.0);
}

// Column grid (binning) and the neighbour capacities sized from it.  Columns
// are about f * cbrt(8 / rho) wide, so that a cluster of 8 atoms is roughly
// cubic at f = 1; (ncx, ncy) > 0 override.  Replaces the compile-time
// CELL_ATOMS / MAXNB presets of the reference (src/configure:89-137).
NbShape nb_shape(const fz_params& par, int n, int ncx, int ncy, double f) {
  NbShape s;
  const double V = par.box[0] * par.box[1] * par.box[2];
  const double rho = n / V;
  const double w = f * std::cbrt(CLUSTER / rho);
  if (ncx <= 0) ncx = std::max(1, (int)std::floor(par.box[0] / w));
  if (ncy <= 0) ncy = std::max(1, (int)std::floor(par.box[1] / w));
  while ((double)n / (ncx * ncy) > 0.5 * MAX_COLUMN_ATOMS) { ncx++; ncy++; }
  s.ncx = ncx; s.ncy = ncy; s.ncol = ncx * ncy;
  s.nzb = std::min(256, std::max(4, (int)std::ceil(par.box[2] / 2.0)));  // ~2 A z bins
  s.clusters_cap = n / CLUSTER + s.ncol + 64;
  const double cw = std::sqrt(par.box[0] / ncx * par.box[1] / ncy);
  const double est = 4.0 / 3.0 * kPi * std::pow(par.pairlist + cw + 1.0, 3) * rho / CLUSTER * 0.5 + 2.0;
  const long long per = std::min<long long>(s.clusters_cap, (long long)std::ceil(1.6 * est));
  s.pairs_cap = (long long)s.clusters_cap * per + 4096;
  return s;
}

static size_t rnd256(size_t b) { return (std::max<size_t>(1, b) + 255) & ~(size_t)255; }

// search waves per I cluster the search buffers are sized for (small systems
// search each I cluster with several waves: more waves in flight)
static int search_parts_of(int clusters_cap) { return clusters_cap < 8192 ? 4 : clusters_cap < 16384 ? 2 : 1; }

// Device bytes of one neighbour block (the allocations of alloc_neighbor_block).
size_t nb_bytes(int ncol, int nzb, int mixed, long long pairs_cap, int clusters_cap) {
  const size_t ns = (size_t)clusters_cap * CLUSTER, cc = clusters_cap;
  const size_t parts = search_parts_of(clusters_cap);
  size_t b = 2 * rnd256(4 * (size_t)ncol) + rnd256(4 * ((size_t)ncol + 1)) + 2 * rnd256(4 * (size_t)ncol * (nzb + 1));
  b += rnd256(4 * ns) + rnd256(ns * (mixed ? 16 : 32)) + rnd256(4 * ns) + rnd256(8 * ns) + rnd256(24 * ns) +
       rnd256(4 * ns) + rnd256(4 * cc) + rnd256(48 * cc) + rnd256(24 * ns);
  b += rnd256(4 * (cc * parts + 1)) + rnd256(4 * cc * parts * SEARCH_TMP_CAP) + rnd256(8 * cc * parts * SEARCH_TMP_CAP);
  b += rnd256(4 * ((cc + 63) / 64));
  b += 2 * rnd256(4 * (size_t)pairs_cap) + rnd256(8 * (size_t)pairs_cap);
  return b;
}

int System::set_column_grid(int ncx, int ncy, double f) {
  SysDev& d = dv;
  const NbShape sh = nb_shape(par, topo.n, ncx, ncy, f);
  free_neighbor();  // the previous neighbour block, if any
  d.ncx = sh.ncx; d.ncy = sh.ncy; d.ncol = sh.ncol;
  // a barostat's box may shrink to NPT_MIN_SCALE of its reference size: the
  // search ranges cover the pair-list distance at that scale
  const double shrink = d.npt ? NPT_MIN_SCALE : 1.0;
  const double bx = d.npt ? d.npt_L0[0] : par.box[0], by = d.npt ? d.npt_L0[1] : par.box[1];
  d.kx_range = (int)std::ceil(par.pairlist / (shrink * bx / sh.ncx)) + 1;
  d.ky_range = (int)std::ceil(par.pairlist / (shrink * by / sh.ncy)) + 1;
  d.colw_x = (unsigned long long)(kTwo64 / sh.ncx);
  d.colw_y = (unsigned long long)(kTwo64 / sh.ncy);
  d.nzb = sh.nzb;
  FZ_TRY(alloc_neighbor(sh.pairs_cap, sh.clusters_cap));
  gen++;
  need_rebuild = true;
  return FZ_OK;
}

int System::alloc_neighbor(long long pairs_cap, int clusters_cap) {
  alloc_nb = true;
  const int rc = alloc_neighbor_block(pairs_cap, clusters_cap);
  alloc_nb = false;
  return rc;
}

int System::alloc_neighbor_block(long long pairs_cap, int clusters_cap) {
  SysDev& d = dv;
  FZ_TRY(dalloc(&d.col_count, d.ncol)); FZ_TRY(dalloc(&d.col_fill, d.ncol)); FZ_TRY(dalloc(&d.col_clu0, d.ncol + 1));
  FZ_TRY(dalloc(&d.zt_hi, (size_t)d.ncol * (d.nzb + 1)));
  FZ_TRY(dalloc(&d.zt_lo, (size_t)d.ncol * (d.nzb + 1)));
  d.n_clusters_cap = clusters_cap;
  d.n_slots_cap = clusters_cap * CLUSTER;
  d.n_pairs_cap = pairs_cap;
  const int ns = d.n_slots_cap;
  FZ_TRY(dalloc(&d.slot_atom, ns));
  FZ_TRY(dalloc((char**)&d.xs, (size_t)ns * (mixed ? 16 : 32)));
  FZ_TRY(dalloc(&d.ts, ns));
  FZ_TRY(dalloc(&d.lp, ns));
  FZ_TRY(dalloc((unsigned long long**)&d.Fs, (size_t)3 * ns));
  FZ_TRY(dalloc(&d.bin_tmp, ns));
  FZ_TRY(dalloc(&d.clu_col, clusters_cap));
  FZ_TRY(dalloc(&d.bb, (size_t)6 * clusters_cap));
  FZ_TRY(dalloc(&d.su, (size_t)3 * ns));
  d.search_parts = search_parts_of(clusters_cap);
  FZ_TRY(dalloc(&d.search_cnt, (size_t)clusters_cap * d.search_parts + 1));
  FZ_TRY(dalloc(&d.search_tmp, (size_t)clusters_cap * d.search_parts * SEARCH_TMP_CAP, false));
  FZ_TRY(dalloc(&d.search_mtmp, (size_t)clusters_cap * d.search_parts * SEARCH_TMP_CAP, false));
  FZ_TRY(dalloc(&d.tile_cnt, (clusters_cap + 63) / 64));
  FZ_TRY(dalloc(&d.pl_i, pairs_cap, false));
  FZ_TRY(dalloc(&d.pl_j, pairs_cap, false));
  FZ_TRY(dalloc(&d.pl_mask, pairs_cap, false));
  return FZ_OK;
}

// Upper bound of a system's device bytes: the allocations of alloc_device
// (exclusion and constraint counts bounded by the term counts) plus the
// neighbour block after one growth of both capacities (the block is replaced,
// not added to, when it grows).
int64_t footprint_bytes(const fz_topology* t, const fz_params* p, int precision) {
  const int mixed = precision == FZ_MIXED ? 1 : 0;
  const size_t n = (size_t)t->n_atoms, T2 = (size_t)t->n_types * t->n_types;
  const size_t nb = t->n_bonds, na = t->n_angles, nd = t->n_dihedrals, ni = t->n_impropers, nc = t->n_constraints;
  const size_t nexcl = 2 * (nb + na + nd) + 1;
  size_t b = 0;
  b += 12 * rnd256(8 * n);                                   // u, build and SHAKE positions, velocities
  b += 3 * rnd256(8 * n) + rnd256(4 * n) + rnd256(4);         // q, minv, mdyn, type, tick
  b += 3 * rnd256(4 * n) + 2 * rnd256(16);                   // com groups, atom_slot, atom_col, counters
  b += rnd256(8 * n) + rnd256(4 * (n + 1)) + rnd256(4 * nexcl);
  b += rnd256(8 * nb) + rnd256(16 * nb) + rnd256(16 * na) + rnd256(32 * na);
  b += rnd256(16 * nd) + rnd256(16 * nd) + rnd256(4 * nd) + rnd256(16 * ni) + rnd256(8 * ni);
  b += rnd256(8 * (nb + na)) + rnd256(8 * nd);               // Ewald-corrected and 1-4 pairs
  b += rnd256(8 * T2) + 2 * rnd256(16 * T2) + rnd256(8 * (size_t)t->n_types);
  b += 3 * rnd256(32 * (nc + 1)) + rnd256(4 * n) + rnd256(8 * n);  // clusters, free atoms, brick atoms
  const size_t K3 = (size_t)p->pme_grid[0] * p->pme_grid[1] * p->pme_grid[2];
  const size_t KH = (size_t)(p->pme_grid[0] / 2 + 1) * p->pme_grid[1] * p->pme_grid[2];
  const size_t KS = (size_t)p->pme_grid[0] + p->pme_grid[1] + p->pme_grid[2];
  size_t nbr = 1;
  for (int k = 0; k < 3; k++) nbr *= (size_t)(p->pme_grid[k] + 7) / 8;
  b += rnd256(8 * K3) + rnd256(KH * (mixed ? 8 : 16)) + rnd256(K3 * (mixed ? 4 : 8)) + rnd256(8 * KH);  // + G(m) table
  b += rnd256(4 * nbr) + rnd256(4 * (nbr + 1)) + rnd256(n * 16 * (mixed ? 4 : 8)) + rnd256(nbr * 1331 * 8);
  b += 3 * rnd256(8 * KS) + rnd256(KS * (mixed ? 8 : 16));
  b += rnd256(8 * A_NSLOT) + rnd256(8 * EV_SPREAD * EV_W) + 2 * rnd256(24 * n) + rnd256(4 * F_NSLOT) + rnd256(16) + 2 * rnd256(24 * n);
  b += rnd256(72 * n + 16);                                  // snapshot
  // the neighbour block after two growth steps of the overflow path (pair-list
  // entries x1.6 + 4096, clusters x1.3 + 64 each): clustered chains can need
  // about twice the pair entries the uniform-density estimate gives
  const NbShape sh = nb_shape(*p, t->n_atoms, p->column_grid[0], p->column_grid[1], 1.0);
  long long pairs = sh.pairs_cap;
  int clus = sh.clusters_cap;
  for (int g = 0; g < 2; g++) {
    pairs = (long long)(pairs * 1.6) + 4096;
    clus = (int)(clus * 1.3) + 64;
  }
  b += nb_bytes(sh.ncol, sh.nzb, mixed, pairs, clus);
  return (int64_t)b;
}

// inverse masses of the integrator: 0 for fixed atoms
int System::upload_mdyn() {
  std::vector<double> m(topo.minv);
  for (int i = 0; i < topo.n; i++)
    if (topo.fixed[i]) m[i] = 0.0;
  dv.n_mobile = topo.n - topo.n_fixed;
  return up(dv.mdyn, m);
}

// Fixed atoms (consfix, src/prmread.cu:699-722,1665-1728; mobility v.w in the
// half kick, src/mdcuda.cu:70-80).  Velocities of fixed atoms are zeroed; the
// constraint clusters are rebuilt without constraints between fixed atoms.
int System::set_fixed(const uint8_t* fixed) {
  const int n = topo.n;
  int nf = 0;
  for (int i = 0; i < n; i++) nf += fixed && fixed[i] ? 1 : 0;
  if (nf == n) return set_error(FZ_ERR_ARG, "every atom is fixed");
  if (nf > 0 && dv.npt) return set_error(FZ_ERR_ARG, "fixed atoms (consfix) cannot join a barostat system");
  for (int i = 0; i < n; i++) topo.fixed[i] = fixed && fixed[i] ? 1 : 0;
  topo.n_fixed = nf;
  FZ_TRY(build_ccl(topo));
  SysDev& d = dv;
  d.n_ccl = (int)topo.ccl_atoms.size();
  // the cluster tables are replaced (no kernel is in flight on this system:
  // calls on one handle do not overlap and every call ends synchronised)
  FZ_HIP(hipDeviceSynchronize());
  dfree(d.ccl_atoms); dfree(d.ccl_cons); dfree(d.ccl_d2); dfree(d.free_atoms);
  d.ccl_atoms = nullptr; d.ccl_cons = nullptr; d.ccl_d2 = nullptr; d.free_atoms = nullptr;
  FZ_TRY(dalloc(&d.ccl_atoms, topo.ccl_atoms.size())); FZ_TRY(up(d.ccl_atoms, topo.ccl_atoms));
  FZ_TRY(dalloc(&d.ccl_cons, topo.ccl_cons.size())); FZ_TRY(up(d.ccl_cons, topo.ccl_cons));
  FZ_TRY(dalloc(&d.ccl_d2, topo.ccl_d2.size())); FZ_TRY(up(d.ccl_d2, topo.ccl_d2));
  d.n_free = (int)topo.free_atoms.size();
  FZ_TRY(dalloc(&d.free_atoms, topo.free_atoms.size())); FZ_TRY(up(d.free_atoms, topo.free_atoms));
  FZ_TRY(upload_mdyn());
  std::vector<double> v((size_t)3 * n);
  FZ_TRY(download_state(nullptr, v.data(), nullptr));
  for (int i = 0; i < n; i++)
    if (topo.fixed[i]) v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0;
  FZ_TRY(upload_velocities(v.data()));
  gen++;
  return FZ_OK;
}

// Harmonic COM restraint (consharm, src/prmread.cu:674-690; src/mdcuda.cu:481-516).
// Andersen barostat (pmass / pref, src/prmread.cu:664-670; piston of
// reduce_virial, src/cell_based_kernels.cu:403-600): the current box becomes the
// reference box L0 (scale 1), the piston starts at rest.  The column search
// ranges are widened so that they hold at the smallest box scale allowed,
// NPT_MIN_SCALE; the cutoff must stay below half of that box.
int System::set_barostat(double pref_atm, double mass) {
  SysDev& d = dv;
  // the box may move from here on: the z pass forms G(m) from the box tables
  if (mass > 0.0 || d.npt) d.pme_g = nullptr;
  if (!(mass > 0.0)) {
    d.npt = 0;
    gen++;  // the descriptor changes (re-uploaded by the next prepare)
    return FZ_OK;
  }
  if (!std::isfinite(pref_atm)) return set_error(FZ_ERR_ARG, "reference pressure must be finite");
  if (topo.n_fixed > 0) return set_error(FZ_ERR_ARG, "the barostat does not take fixed atoms (consfix)");
  for (int k = 0; k < 3; k++)
    if (!(par.cutoff < 0.5 * NPT_MIN_SCALE * d.L[k]))
      return set_error(FZ_ERR_ARG, "barostat: cutoff %.3f must stay below half the box at %.2f of its size (%.3f)",
                       par.cutoff, NPT_MIN_SCALE, 0.5 * NPT_MIN_SCALE * d.L[k]);
  if (!d.nst) FZ_TRY(dalloc((char**)&d.nst, 2 * sizeof(NptState)));
  d.npt = 1;
  d.npt_pref = pref_atm * kPU;
  d.npt_mass = mass;
  for (int k = 0; k < 3; k++) d.npt_L0[k] = d.L[k];
  d.npt_V0 = d.L[0] * d.L[1] * d.L[2];
  d.kx_range = (int)std::ceil(par.pairlist / (NPT_MIN_SCALE * d.npt_L0[0] / d.ncx)) + 1;
  d.ky_range = (int)std::ceil(par.pairlist / (NPT_MIN_SCALE * d.npt_L0[1] / d.ncy)) + 1;
  NptState st{};
  st.V = d.npt_V0;
  st.ell = st.ell_b = 1.0;
  st.nbuild = flags_host[F_NBUILD];
  npt_host = st;
  FZ_HIP(hipMemcpy(d.nst, &st, sizeof st, hipMemcpyHostToDevice));
  npt_box_fields(d, 1.0, 1.0);
  need_rebuild = true;
  gen++;
  return FZ_OK;
}

void System::npt_sync_host() {
  if (!dv.npt) return;
  npt_box_fields(dv, npt_host.ell, npt_host.ell_b);
  for (int k = 0; k < 3; k++) par.box[k] = dv.L[k];
}

int System::set_com_restraint(const int32_t* group, const double dist[3], double fc) {
  const int n = topo.n;
  SysDev& d = dv;
  if (!group || fc == 0.0) {
    d.com_on = 0;
    gen++;
    return FZ_OK;
  }
  if (!std::isfinite(fc)) return set_error(FZ_ERR_ARG, "restraint force constant must be finite");
  double M[2] = {0.0, 0.0};
  std::vector<int> g(n);
  for (int i = 0; i < n; i++) {
    if (group[i] < -1 || group[i] > 1) return set_error(FZ_ERR_ARG, "restraint group of atom %d must be -1, 0 or 1", i);
    g[i] = group[i];
    if (g[i] >= 0) M[g[i]] += topo.mass[i];
  }
  if (!(M[0] > 0.0 && M[1] > 0.0)) return set_error(FZ_ERR_ARG, "both restraint groups need atoms");
  FZ_TRY(up(d.com_grp, g));
  d.com_on = 1;
  for (int k = 0; k < 3; k++) d.com_d[k] = dist[k];
  d.com_fc = fc;
  d.com_im[0] = 1.0 / M[0];
  d.com_im[1] = 1.0 / M[1];
  gen++;
  return FZ_OK;
}

int System::upload_fixed(const uint64_t* u) {
  const int n = topo.n;
  std::vector<unsigned long long> c(n);
  unsigned long long* dst[3] = {dv.ux, dv.uy, dv.uz};
  for (int k = 0; k < 3; k++) {
    for (int i = 0; i < n; i++) c[i] = u[3 * i + k];
    FZ_HIP(hipMemcpy(dst[k], c.data(), n * sizeof(unsigned long long), hipMemcpyHostToDevice));
  }
  need_rebuild = true;
  min_step = 0.0;
  return FZ_OK;
}

int System::upload_positions(const double* xyz) {
  const int n = topo.n;
  std::vector<uint64_t> u((size_t)3 * n);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double t = xyz[3 * i + k] / dv.L[k];
      t -= std::floor(t);
      if (!(t < 1.0)) t = 0.0;
      u[3 * i + k] = (uint64_t)(t * kTwo64);
    }
  return upload_fixed(u.data());
}

int System::upload_velocities(const double* v) {
  const int n = topo.n;
  std::vector<double> c(n, 0.0);
  double* dst[3] = {dv.vx, dv.vy, dv.vz};
  for (int k = 0; k < 3; k++) {
    if (v)
      for (int i = 0; i < n; i++) c[i] = v[3 * i + k];
    FZ_HIP(hipMemcpy(dst[k], c.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
  return FZ_OK;
}

int System::download_state(double* xyz, double* v, uint64_t* u) {
  const int n = topo.n;
  if (xyz || u) {
    std::vector<unsigned long long> c(n);
    unsigned long long* src[3] = {dv.ux, dv.uy, dv.uz};
    for (int k = 0; k < 3; k++) {
      FZ_HIP(hipMemcpy(c.data(), src[k], n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++) {
        if (u) u[3 * i + k] = c[i];
        if (xyz) {
          double x = (double)c[i] * dv.sc[k];
          if (x >= 0.5 * dv.L[k]) x -= dv.L[k];
          xyz[3 * i + k] = x;
        }
      }
    }
  }
  if (v) {
    std::vector<double> c(n);
    double* src[3] = {dv.vx, dv.vy, dv.vz};
    for (int k = 0; k < 3; k++) {
      FZ_HIP(hipMemcpy(c.data(), src[k], n * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++) v[3 * i + k] = c[i];
    }
  }
  return FZ_OK;
}

// energy / virial accumulators are int64 fixed point at 2^28 per kcal/mol
static inline double fx(unsigned long long v) { return (double)(long long)v / 268435456.0; }

int System::energies(fz_energies* e) {
  memset(e, 0, sizeof *e);
  const unsigned long long* a = acc_host;
  e->step = step;
  e->bond = fx(a[A_BOND]);
  e->angle = fx(a[A_ANGLE]);
  e->urey_bradley = fx(a[A_UB]);
  e->improper = fx(a[A_IMPR]);
  e->dihedral = fx(a[A_DIHED]);
  e->vdw = fx(a[A_VDW]);
  e->elec = fx(a[A_ELEC]);
  e->pme = fx(a[A_PME]);
  e->ewald_excl = fx(a[A_EWEXCL]);
  e->ewald_self = self_energy;
  e->potential = e->bond + e->angle + e->urey_bradley + e->improper + e->dihedral + e->vdw + e->elec + e->pme +
                 e->ewald_excl + e->ewald_self;
  e->kinetic = fx(a[A_KIN]);
  // 3 N_mobile - 6 - N_constraints (src/evalprops.cu:221-225)
  const double dof = 3.0 * (topo.n - topo.n_fixed) - 6.0 - topo.n_cons_active;
  e->temperature = dof > 0 ? 2.0 * e->kinetic / (kRC * dof) : 0.0;
  for (int k = 0; k < 9; k++) e->virial[k] = fx(a[A_VIR + k]);
  e->pairs_within_cutoff = (int64_t)a[A_NPAIR];
  e->pair_tiles = (int64_t)a[A_NTILE];
  e->list_builds = flags_host[F_NBUILD];
  e->volume = dv.L[0] * dv.L[1] * dv.L[2];
  // (2 KE + tr W) / 3V in atm (the reference's pressure, src/cell_based_kernels.cu:586-589)
  e->pressure = (2.0 * e->kinetic + e->virial[0] + e->virial[4] + e->virial[8]) / (3.0 * e->volume) / kPU;
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// Engine

Engine::Engine(int device, std::vector<System*> sys) : device_(device), sys_(std::move(sys)) {}

Engine::~Engine() {
  drop_graphs();
  for (int* p : d_pre_) (void)hipFree(p);
  if (d_sys_) (void)hipFree(d_sys_);
  if (L_.chunk_pre) (void)hipFree(L_.chunk_pre);
  if (L_.chunk_pre_a) (void)hipFree(L_.chunk_pre_a);
  for (int k = 0; k < NSIDE; k++) {
    if (ev_join_[k]) (void)hipEventDestroy(ev_join_[k]);
    if (side_[k]) (void)hipStreamDestroy(side_[k]);
  }
  if (ev_early_) (void)hipEventDestroy(ev_early_);
  if (ev_split_) (void)hipEventDestroy(ev_split_);
  if (L_.stream) (void)hipStreamDestroy(L_.stream);
}

void Engine::drop_graphs() {
  if (g_plain_) (void)hipGraphExecDestroy(g_plain_);
  if (g_ev_) (void)hipGraphExecDestroy(g_ev_);
  if (g_npt_) (void)hipGraphExecDestroy(g_npt_);
  if (g_plain_k_) (void)hipGraphExecDestroy(g_plain_k_);
  if (g_npt_k_) (void)hipGraphExecDestroy(g_npt_k_);
  g_plain_ = g_ev_ = g_npt_ = g_plain_k_ = g_npt_k_ = nullptr;
}

int Engine::make_work(Work& w, const std::vector<int>& blocks) {
  // block prefixes per system, then the system of every block (block_sys)
  std::vector<int> pre(blocks.size() + 1, 0);
  for (size_t s = 0; s < blocks.size(); s++) pre[s + 1] = pre[s] + blocks[s];
  const int total = pre.back();
  for (size_t s = 0; s < blocks.size(); s++)
    for (int k = 0; k < blocks[s]; k++) pre.push_back((int)s);
  int* d = nullptr;
  FZ_HIP(hipMalloc(&d, pre.size() * sizeof(int)));
  d_pre_.push_back(d);
  FZ_HIP(hipMemcpy(d, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
  w.pre = d;
  w.total = total;
  return FZ_OK;
}

int Engine::upload_desc() {
  std::vector<SysDev> h(sys_.size());
  for (size_t s = 0; s < sys_.size(); s++) h[s] = sys_[s]->dv;
  FZ_HIP(hipMemcpy(d_sys_, h.data(), h.size() * sizeof(SysDev), hipMemcpyHostToDevice));
  return FZ_OK;
}

int Engine::prepare() {
  FZ_HIP(hipSetDevice(device_));
  drop_graphs();
  for (int* p : d_pre_) (void)hipFree(p);
  d_pre_.clear();
  const int ns = (int)sys_.size();
  if (ns == 0) return set_error(FZ_ERR_ARG, "empty system set");
  if (!L_.stream) FZ_HIP(hipStreamCreateWithFlags(&L_.stream, hipStreamNonBlocking));
  for (int k = 0; k < NSIDE; k++) {
    if (!side_[k]) FZ_HIP(hipStreamCreateWithFlags(&side_[k], hipStreamNonBlocking));
    if (!ev_join_[k]) FZ_HIP(hipEventCreateWithFlags(&ev_join_[k], hipEventDisableTiming));
  }
  if (!ev_early_) FZ_HIP(hipEventCreateWithFlags(&ev_early_, hipEventDisableTiming));
  if (!ev_split_) FZ_HIP(hipEventCreateWithFlags(&ev_split_, hipEventDisableTiming));
  if (!d_sys_) FZ_HIP(hipMalloc(&d_sys_, ns * sizeof(SysDev)));
  for (int** cp : {&L_.chunk_pre, &L_.chunk_pre_a}) {
    if (*cp) continue;
    FZ_HIP(hipMalloc(cp, (ns + 1) * sizeof(int)));
    FZ_HIP(hipMemset(*cp, 0, (ns + 1) * sizeof(int)));
  }
  FZ_TRY(upload_desc());
  L_.d_sys = d_sys_;
  L_.nsys = ns;
  L_.mixed = sys_[0]->mixed;
  for (auto* s : sys_)
    if (s->mixed != L_.mixed) return set_error(FZ_ERR_ARG, "a batch must use one precision");
  L_.npt = sys_[0]->dv.npt;
  for (auto* s : sys_)
    if (s->dv.npt != L_.npt) return set_error(FZ_ERR_ARG, "a batch must be all barostat or all fixed-box systems");
  L_.lj_sep = 1;
  L_.com_any = 0;
  for (auto* s : sys_) {
    L_.lj_sep &= s->dv.lj_sep;
    L_.com_any |= s->dv.com_on;
  }
  std::vector<int> atoms, atoms64, slots, cols, clus, srch, terms, ccl, stp, fx_, fy_, fz_, bricks, planes;
  // Waves per I cluster in the pair search: up to each system's search_parts
  // when the launch holds few clusters (latency bound: too few waves in
  // flight), one when it holds many.
  long long clusters_total = 0;
  for (auto* s : sys_) clusters_total += s->dv.n_clusters_cap;
  L_.search_cap = clusters_total <= 16384 ? 4 : 1;
  int kmax = 4;
  for (auto* s : sys_) kmax = std::max({kmax, s->dv.K[0], s->dv.K[1], s->dv.K[2]});
  const int csz = L_.mixed ? 8 : 16;
  int NL = tuned_nl_ > 0 ? tuned_nl_ : 16;
  while (NL > 1 && (size_t)(NL * 2 + 1) * kmax * csz > 49152) NL >>= 1;
  L_.fft_nl = NL;
  L_.fft_lds = (NL * 2 + 1) * kmax * csz;  // two line buffers + the axis' twiddles
  for (auto* s : sys_) {
    const SysDev& d = s->dv;
    atoms.push_back((d.n_atoms + 255) / 256);
    atoms64.push_back((d.n_atoms + 63) / 64);
    slots.push_back((d.n_slots_cap + 255) / 256);
    cols.push_back(d.ncol);
    clus.push_back((d.n_clusters_cap + 3) / 4);
    srch.push_back((d.n_clusters_cap * search_split(d.search_parts, L_.search_cap) + 3) / 4);
    const int nt = d.n_bonds + d.n_angles + d.n_diheds + d.n_imprs + d.n_ew + d.n_p14;
    terms.push_back((nt + 255) / 256);
    ccl.push_back((d.n_ccl + 127) / 128);
    stp.push_back((d.n_ccl + d.n_free + 127) / 128);
    const int ng = (d.Kxh + NL - 1) / NL;
    fx_.push_back((d.K[1] * d.K[2] + NL - 1) / NL);
    fy_.push_back(d.K[2] * ng);
    fz_.push_back(d.K[1] * ng);
    bricks.push_back(d.n_bricks);
    planes.push_back(d.K[2]);
  }
  FZ_TRY(make_work(L_.w_atoms, atoms));
  FZ_TRY(make_work(L_.w_atoms64, atoms64));
  FZ_TRY(make_work(L_.w_slots, slots));
  FZ_TRY(make_work(L_.w_cols, cols));
  FZ_TRY(make_work(L_.w_clus, clus));
  FZ_TRY(make_work(L_.w_search, srch));
  FZ_TRY(make_work(L_.w_terms, terms));
  FZ_TRY(make_work(L_.w_ccl, ccl));
  FZ_TRY(make_work(L_.w_step, stp));
  FZ_TRY(make_work(L_.w_fftx, fx_));
  FZ_TRY(make_work(L_.w_ffty, fy_));
  FZ_TRY(make_work(L_.w_fftz, fz_));
  FZ_TRY(make_work(L_.w_bricks, bricks));
  FZ_TRY(make_work(L_.w_planes, planes));
  // Default persistent pair grid: about ten list chunks per block (a list holds
  // ~19 cluster-pair entries per atom at water density, 32 entries per chunk);
  // measured on the 17K slab, 1,024 blocks beat 640 and 2,048 (the bonded and
  // PME branches run beside the pair kernel).
  long long atoms_total = 0;
  for (auto* s : sys_) atoms_total += s->dv.n_atoms;
  const int grid0 = (int)std::min<long long>(4096, std::max<long long>(256, (atoms_total / 16 + 63) / 64 * 64));
  L_.pair_grid = tuned_grid_ > 0 ? tuned_grid_ : grid0;
  // Spread by bricks from 64K atoms per launch; below, the single atomic kernel
  // is cheaper than the brick pipeline's five launches.  Same grid either way.
  L_.spread_bricks = atoms_total >= 65536 ? 1 : 0;
  // Systems of up to PLANE_MAX_ATOMS atoms: spread and x, y transforms by the
  // plane kernels (two launches and no grid atomics instead of four launches).
  L_.plane_lds = 0;
  L_.pme_planes = 1;
  for (auto* s : sys_) {
    L_.plane_lds = std::max(L_.plane_lds, plane_lds_bytes(s->dv.K[0], s->dv.K[1], csz));
    if (s->dv.n_atoms > PLANE_MAX_ATOMS) L_.pme_planes = 0;
  }
  if (tuned_planes_ == 1) L_.pme_planes = 1;  // asked for: any atom count
  if (L_.plane_lds > PLANE_MAX_LDS || tuned_planes_ == 0) L_.pme_planes = 0;
  prepared_ = true;
  gens_.clear();
  for (auto* s : sys_) gens_.push_back(s->gen);
  return FZ_OK;
}

int Engine::ensure() {
  bool stale = !prepared_ || gens_.size() != sys_.size();
  for (size_t k = 0; !stale && k < sys_.size(); k++) stale = gens_[k] != sys_[k]->gen;
  if (stale) FZ_TRY(prepare());
  // The atomic spread adds into a zero grid, which its x pass leaves behind;
  // the brick spread overwrites the grid and leaves its last spread there.  A
  // member whose grid was last spread by bricks (in another engine: a batch
  // over 64K atoms) is cleared before this engine spreads atomically.
  for (auto* s : sys_) {
    if (L_.pme_planes) continue;  // the plane kernels neither read nor clear the grid accumulator
    if (!L_.spread_bricks && s->grid_mode != 0 && s->dv.Qacc) {
      const size_t K3 = (size_t)s->dv.K[0] * s->dv.K[1] * s->dv.K[2];
      FZ_HIP(hipMemsetAsync(s->dv.Qacc, 0, K3 * 8, L_.stream));
    }
    s->grid_mode = L_.spread_bricks;
  }
  return FZ_OK;
}

void Engine::record_forces(bool ev, bool lite, bool check) {
  // Fork: bonded and PME beside the neighbour build and the pair kernel
  // (errors of these calls surface at the stream's next synchronisation or at
  // capture end).  Their forces go to atom-ordered sums, and the charge spread
  // and gather visit atoms, so neither branch waits for the build's slot
  // layout: both start with the build.
  Launch side[NSIDE] = {L_, L_, L_};
  for (int k = 0; k < NSIDE; k++) side[k].stream = side_[k];
  // Batch dynamics: the pair kernel runs twice.  On the third side stream, at
  // once, over the members whose lists stay valid this step (no rebuild
  // pending); on the main stream, after the build, over the members it
  // rebuilt.  The build of a few members then overlaps the pair work of the
  // others.  (For one system the extra launches and cross-stream wait cost
  // more than the build's launches they hide, at 17K and at 230K atoms.)
  const bool split = !check && L_.nsys > 1;
  const int nside = split ? 3 : 2;
  const int pev = ev ? (lite ? 1 : 2) : 0;
  (void)hipEventRecord(ev_early_, L_.stream);
  for (int k = 0; k < nside; k++) (void)hipStreamWaitEvent(side_[k], ev_early_, 0);
  launch_pme_part(side[1], 0, ev);
  launch_bonded(side[0], ev);
  launch_pme_part(side[1], 1, ev);
  launch_pme_part(side[1], 2, ev);
  if (split) {
    launch_chunk_split(side[2], false);
    (void)hipEventRecord(ev_split_, side_[2]);
    launch_pair(side[2], pev, L_.chunk_pre_a, false);
  }
  launch_nblist(L_, check);
  if (check) {
    // stand-alone evaluation: slot copies and every system's work items
    launch_gather_sorted(L_);
    launch_pair(L_, pev);
  } else if (split) {
    // the split above reads the rebuild decisions before this one spends them
    (void)hipStreamWaitEvent(L_.stream, ev_split_, 0);
    launch_chunk_split(L_, true);
    launch_pair(L_, pev, L_.chunk_pre, false);
  } else {
    // one system in dynamics: the slot copies are kept current by k_step_first
    // and the build, which also lays out the work items
    launch_pair(L_, pev, L_.chunk_pre, false, true);
  }
  for (int k = 0; k < nside; k++) {
    (void)hipEventRecord(ev_join_[k], side_[k]);
    (void)hipStreamWaitEvent(L_.stream, ev_join_[k], 0);
  }
  if (!check && pev) launch_ev_fold(L_);
  if (L_.com_any) launch_restraint(L_);
}

// A run of n steps is: first half (eager), then n - 1 replays of [forces,
// second half of this step fused with the first half of the next], then one
// replay of [forces with energies, second half, kinetic energy].  Every
// replay thus holds one integrator kernel between force evaluations; the
// arithmetic is that of separate halves.  The step's pair energies are summed
// in FP64 from the FP32 pair terms (the mixed form's arithmetic; a stand-alone
// evaluation recomputes them in FP64, which costs the 17K slab's pair kernel
// seven times its force-only time).
void Engine::record_step(bool ev) {
  record_forces(ev, true, false);
  if (!ev) {
    launch_step_first(L_, true);
  } else {
    launch_step_second(L_);
    launch_kinetic(L_);
  }
}

// One barostat step (see the NPT kernels in fz_integrate.hip): piston and box
// halves around the drift, SHAKE at the new box, forces with the virial every
// step (pair energies from the FP32 terms), the clusters' internal virial and
// the piston half kick, kick and RATTLE, kinetic energy.
void Engine::record_npt_step() {
  launch_npt_a(L_);
  launch_npt_b(L_);
  launch_npt_c(L_);
  launch_npt_d(L_);
  record_forces(true, true, false);
  launch_npt_intra(L_);
  launch_npt_g(L_);
  launch_step_second(L_);
  launch_kinetic(L_);
}

int Engine::capture_npt(hipGraphExec_t* out, int nsteps) {
  hipGraph_t g = nullptr;
  FZ_HIP(hipStreamBeginCapture(L_.stream, hipStreamCaptureModeRelaxed));
  for (int k = 0; k < nsteps; k++) record_npt_step();
  FZ_HIP(hipStreamEndCapture(L_.stream, &g));
  hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_error(FZ_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  return FZ_OK;
}

// n barostat steps, one replay each; failure handling as in steps()
int Engine::steps_npt(int n) {
  for (auto* s : sys_)
    if (s->need_rebuild) FZ_TRY(forces(true));
  for (int attempt = 0; attempt < 4; attempt++) {
    FZ_TRY(snapshot(true));
    if (!g_npt_) FZ_TRY(capture_npt(&g_npt_));
    if (!g_npt_k_ && (sys_.size() == 1 || n >= GRAPH_STEPS)) FZ_TRY(capture_npt(&g_npt_k_, GRAPH_STEPS));
    launch_gather_sorted(L_);
    int k = 0;
    for (; k + GRAPH_STEPS <= n; k += GRAPH_STEPS) FZ_HIP(hipGraphLaunch(g_npt_k_, L_.stream));
    for (; k < n; k++) FZ_HIP(hipGraphLaunch(g_npt_, L_.stream));
    FZ_TRY(read_status());
    std::vector<int> grow(sys_.size(), 0);
    int overflow = 0, fatal = -1, shake = -1, box = -1;
    for (size_t k = 0; k < sys_.size(); k++) {
      const int ov = sys_[k]->flags_host[F_OVERFLOW];
      grow[k] = ov & ~OV_BOX;
      overflow |= grow[k];
      if ((ov & (OV_COLUMN | OV_SEARCH)) && fatal < 0) fatal = (int)k;
      if ((ov & OV_BOX) && box < 0) box = (int)k;
      if (sys_[k]->flags_host[F_SHAKE_FAIL] && shake < 0) shake = (int)k;
    }
    if (!overflow && shake < 0 && box < 0) {
      for (auto* s : sys_) s->step += n;
      return FZ_OK;
    }
    if (fatal >= 0 || box >= 0 || (shake >= 0 && !overflow)) {
      FZ_TRY(restore());
      if (fatal >= 0)
        return set_error(FZ_ERR_CAPACITY, "neighbour search capacity exceeded in system %d of the set (column "
                         "occupancy or more than %d J clusters in a search part); use a finer column grid",
                         fatal, SEARCH_TMP_CAP);
      if (box >= 0)
        return set_error(FZ_ERR_STATE, "barostat: the box of system %d shrank below %.2f of its reference size",
                         box, NPT_MIN_SCALE);
      return set_error(FZ_ERR_SHAKE, "SHAKE did not converge (system %d of the set)", shake);
    }
    FZ_TRY(restore(&grow));
  }
  return set_error(FZ_ERR_CAPACITY, "neighbour capacity could not be grown");
}

int Engine::capture(bool ev, hipGraphExec_t* out, int nsteps) {
  hipGraph_t g = nullptr;
  FZ_HIP(hipStreamBeginCapture(L_.stream, hipStreamCaptureModeRelaxed));
  for (int k = 0; k < nsteps; k++) record_step(ev);
  FZ_HIP(hipStreamEndCapture(L_.stream, &g));
  hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_error(FZ_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  return FZ_OK;
}

int Engine::read_status() {
  for (auto* s : sys_) {
    FZ_HIP(hipMemcpyAsync(s->flags_host, s->dv.flags, sizeof(s->flags_host), hipMemcpyDeviceToHost, L_.stream));
    FZ_HIP(hipMemcpyAsync(s->acc_host, s->dv.acc, sizeof(s->acc_host), hipMemcpyDeviceToHost, L_.stream));
    if (s->dv.npt)
      FZ_HIP(hipMemcpyAsync(&s->npt_host, s->dv.nst, sizeof(NptState), hipMemcpyDeviceToHost, L_.stream));
  }
  FZ_HIP(hipStreamSynchronize(L_.stream));
  for (auto* s : sys_) s->npt_sync_host();
  return FZ_OK;
}

int Engine::npt_reset_virial() {
  if (!L_.npt) return FZ_OK;
  launch_npt_wtr(L_);
  FZ_HIP(hipGetLastError());
  return read_status();
}

int Engine::sync() {
  FZ_HIP(hipStreamSynchronize(L_.stream));
  return FZ_OK;
}

// Neighbour arrays of member si at new capacities; the list they held is gone
// (the member rebuilds at its next evaluation).
int Engine::realloc_neighbor(int si, long long pairs, int clus) {
  System* s = sys_[si];
  SysDev& d = s->dv;
  FZ_HIP(hipStreamSynchronize(L_.stream));
  s->free_neighbor();
  FZ_TRY(s->alloc_neighbor(pairs, clus));
  s->gen++;
  FZ_HIP(hipMemset(d.flags + F_OVERFLOW, 0, sizeof(int)));
  FZ_HIP(hipMemset(d.col_count, 0, d.ncol * sizeof(int)));
  FZ_HIP(hipMemset(d.col_fill, 0, d.ncol * sizeof(int)));
  FZ_HIP(hipMemset(d.tick, 0, sizeof(unsigned)));
  s->need_rebuild = true;
  return FZ_OK;
}

int Engine::grow_and_reset(int si, int bits) {
  const SysDev& d = sys_[si]->dv;
  long long pairs = d.n_pairs_cap;
  int clus = d.n_clusters_cap;
  if (bits & (OV_COLUMN | OV_SEARCH))
    return set_error(FZ_ERR_CAPACITY, "neighbour search capacity exceeded (column occupancy or more than %d J clusters "
                     "in a search part); use a finer column grid", SEARCH_TMP_CAP);
  if (bits & OV_PAIRS) pairs = (long long)(pairs * 1.6) + 4096;
  if (bits & OV_SLOTS) clus = (int)(clus * 1.3) + 64;
  return realloc_neighbor(si, pairs, clus);
}

static unsigned long long* snap_part(System* s, int k) { return s->snap + (size_t)k * s->topo.n; }

int Engine::snapshot(bool clear_shake) {
  // one launch over every member (positions, velocities, build positions,
  // flags); clear_shake also clears the SHAKE-failure flag first, so that it
  // reports the steps after the snapshot only
  launch_snapshot(L_, clear_shake);
  FZ_HIP(hipGetLastError());
  for (auto* s : sys_) {
    s->snap_step = s->step;
    s->snap_min_step = s->min_step;
  }
  return FZ_OK;
}

// Every member returns to its snapshot.  Its list is rebuilt at the snapshot's
// build positions, which reproduces the list it held (the build depends only
// on those positions and the column grid, not on the capacities), then the
// positions go back to the snapshot's and the forces are evaluated there (no
// rebuild: no atom was past half the skin from the build positions, or the
// list would have been rebuilt then).  Velocities, flags (list-build count
// included) and the host counters are the snapshot's, so a replay from here
// repeats what the interrupted call computed.
int Engine::restore(const std::vector<int>* grow) {
  for (size_t k = 0; k < sys_.size(); k++) {
    System* s = sys_[k];
    const size_t nb = (size_t)s->topo.n * 8;
    void* v[3] = {s->dv.vx, s->dv.vy, s->dv.vz};
    void* u[3] = {s->dv.ux, s->dv.uy, s->dv.uz};
    for (int c = 0; c < 3; c++) {
      FZ_HIP(hipMemcpyAsync(v[c], snap_part(s, 3 + c), nb, hipMemcpyDeviceToDevice, L_.stream));
      FZ_HIP(hipMemcpyAsync(u[c], snap_part(s, 6 + c), nb, hipMemcpyDeviceToDevice, L_.stream));
    }
    FZ_HIP(hipMemcpyAsync(s->dv.flags, snap_part(s, 9), F_NSLOT * sizeof(int), hipMemcpyDeviceToDevice, L_.stream));
    if (s->dv.npt)
      FZ_HIP(hipMemcpyAsync(s->dv.nst, s->dv.nst + 1, sizeof(NptState), hipMemcpyDeviceToDevice, L_.stream));
    if (grow && (*grow)[k]) FZ_TRY(grow_and_reset((int)k, (*grow)[k]));
    s->need_rebuild = true;
    s->step = s->snap_step;
    s->min_step = s->snap_min_step;
  }
  FZ_TRY(ensure());
  // barostat: the list is rebuilt at the box of its build, the forces are then
  // evaluated at the snapshot's box
  if (L_.npt) launch_npt_apply(L_, true);
  FZ_TRY(forces(true));
  for (auto* s : sys_) {
    const size_t nb = (size_t)s->topo.n * 8;
    void* u[3] = {s->dv.ux, s->dv.uy, s->dv.uz};
    for (int c = 0; c < 3; c++)
      FZ_HIP(hipMemcpyAsync(u[c], snap_part(s, c), nb, hipMemcpyDeviceToDevice, L_.stream));
  }
  // (the evaluation above moved the state's build mark: the snapshot's again)
  for (auto* s : sys_)
    if (s->dv.npt)
      FZ_HIP(hipMemcpyAsync(s->dv.nst, s->dv.nst + 1, sizeof(NptState), hipMemcpyDeviceToDevice, L_.stream));
  if (L_.npt) launch_npt_apply(L_, false);
  FZ_TRY(forces(true));
  // the list-build count as at the snapshot (the two evaluations above rebuilt
  // a list the trajectory already had); the other flags are clear there too;
  // the barostat state as at the snapshot (the evaluations moved its build mark)
  for (auto* s : sys_) {
    FZ_HIP(hipMemcpyAsync(s->dv.flags, snap_part(s, 9), F_NSLOT * sizeof(int), hipMemcpyDeviceToDevice, L_.stream));
    if (s->dv.npt)
      FZ_HIP(hipMemcpyAsync(s->dv.nst, s->dv.nst + 1, sizeof(NptState), hipMemcpyDeviceToDevice, L_.stream));
  }
  return read_status();
}

int Engine::set_capacity(int si, long long pairs_cap, int clusters_cap) {
  if (si < 0 || si >= (int)sys_.size()) return set_error(FZ_ERR_ARG, "no member %d", si);
  System* s = sys_[si];
  if (pairs_cap < 0 || clusters_cap < 0) return set_error(FZ_ERR_ARG, "capacities must not be negative");
  if (clusters_cap > 0 && clusters_cap < s->topo.n / CLUSTER + 1)
    return set_error(FZ_ERR_ARG, "cluster capacity %d cannot hold %d atoms", clusters_cap, s->topo.n);
  FZ_TRY(ensure());
  for (auto* m : sys_)
    if (m->need_rebuild) FZ_TRY(forces(true));
  FZ_TRY(snapshot());
  FZ_TRY(realloc_neighbor(si, pairs_cap > 0 ? pairs_cap : s->dv.n_pairs_cap,
                          clusters_cap > 0 ? clusters_cap : s->dv.n_clusters_cap));
  return restore();
}

int Engine::forces(bool ev, bool lite) {
  FZ_TRY(ensure());
  for (int attempt = 0; attempt < 4; attempt++) {
    for (auto* s : sys_) {
      if (s->need_rebuild) {
        int one = 1;
        FZ_HIP(hipMemcpyAsync(s->dv.flags + F_REBUILD, &one, sizeof(int), hipMemcpyHostToDevice, L_.stream));
        FZ_HIP(hipStreamSynchronize(L_.stream));
      }
    }
    launch_zero_acc(L_);
    launch_zero_forces(L_);
    record_forces(ev, lite);
    launch_kinetic(L_);
    if (L_.npt) launch_npt_mark(L_);
    FZ_HIP(hipGetLastError());
    FZ_TRY(read_status());
    bool again = false;
    for (size_t k = 0; k < sys_.size(); k++) {
      System* s = sys_[k];
      if (s->flags_host[F_OVERFLOW]) {
        FZ_TRY(grow_and_reset((int)k, s->flags_host[F_OVERFLOW]));
        again = true;
      } else {
        s->need_rebuild = false;
      }
    }
    if (!again) return FZ_OK;
    FZ_TRY(prepare());
  }
  return set_error(FZ_ERR_CAPACITY, "neighbour capacity could not be grown");
}

int Engine::steps(int n) {
  if (n <= 0) return FZ_OK;
  FZ_TRY(ensure());
  if (L_.npt) return steps_npt(n);
  for (auto* s : sys_)
    if (s->need_rebuild) FZ_TRY(forces(true));
  for (int attempt = 0; attempt < 4; attempt++) {
    FZ_TRY(snapshot(true));  // the SHAKE flag reports this attempt's steps only
    if (!g_plain_) FZ_TRY(capture(false, &g_plain_));
    if (!g_ev_) FZ_TRY(capture(true, &g_ev_));
    // a single system captures it with the others (so that no later call pays
    // for its capture); a batch on its first call long enough to use it
    if (!g_plain_k_ && (sys_.size() == 1 || n - 1 >= GRAPH_STEPS)) FZ_TRY(capture(false, &g_plain_k_, GRAPH_STEPS));
    // The step graph has no slot gather: it relies on the slot copies and
    // work items of the last list build or force evaluation, which may belong
    // to another engine over the same systems, so lay them out for this one
    // first.
    launch_gather_sorted(L_);
    launch_step_first(L_);
    int k = 0;
    for (; k + GRAPH_STEPS <= n - 1; k += GRAPH_STEPS) FZ_HIP(hipGraphLaunch(g_plain_k_, L_.stream));
    for (; k < n - 1; k++) FZ_HIP(hipGraphLaunch(g_plain_, L_.stream));
    FZ_HIP(hipGraphLaunch(g_ev_, L_.stream));
    FZ_TRY(read_status());
    // classify before touching any state
    std::vector<int> grow(sys_.size(), 0);
    int overflow = 0, fatal = -1, shake = -1;
    for (size_t k = 0; k < sys_.size(); k++) {
      const int ov = sys_[k]->flags_host[F_OVERFLOW];
      grow[k] = ov;
      overflow |= ov;
      if ((ov & (OV_COLUMN | OV_SEARCH)) && fatal < 0) fatal = (int)k;
      if (sys_[k]->flags_host[F_SHAKE_FAIL] && shake < 0) shake = (int)k;
    }
    if (!overflow && shake < 0) {
      for (auto* s : sys_) s->step += n;
      return FZ_OK;
    }
    if (fatal >= 0 || (shake >= 0 && !overflow)) {
      // an error the replay cannot cure: every member back to the state before
      // the call, then report it
      FZ_TRY(restore());
      if (fatal >= 0)
        return set_error(FZ_ERR_CAPACITY, "neighbour search capacity exceeded in system %d of the set (column "
                         "occupancy or more than %d J clusters in a search part); use a finer column grid",
                         fatal, SEARCH_TMP_CAP);
      return set_error(FZ_ERR_SHAKE, "SHAKE did not converge (system %d of the set)", shake);
    }
    // A capacity overflow (a SHAKE failure after it may be its consequence:
    // the steps ran without the pairs that did not fit): every member returns
    // to the snapshot, the overflowed ones with larger capacities, and the
    // steps are replayed.
    FZ_TRY(restore(&grow));
  }
  return set_error(FZ_ERR_CAPACITY, "neighbour capacity could not be grown");
}

void launch_min_move(const Launch& L);
void launch_min_restore(const Launch& L, const int* d_reject);
void launch_zero_vel(const Launch& L);

// Steepest descent with an adaptive step (replaces MinimizeStep's SD branch,
// src/dynamics.cu:491-620, src/mdcuda.cu:245): the atom with the largest force
// moves by the current step; an energy decrease is accepted and grows the step
// by 1.2, an increase is undone and halves it (the reference reverts and
// switches method on an increase, src/dynamics.cu:597-618).
int Engine::minimize_sd(int nsteps, double max_step) {
  const int ns = (int)sys_.size();
  std::vector<double> step(ns, max_step), e_old(ns);
  for (int s = 0; s < ns; s++)
    if (sys_[s]->min_step > 0.0) step[s] = std::min(sys_[s]->min_step, max_step);
  std::vector<int> reject(ns, 0);
  int* d_reject = nullptr;
  FZ_HIP(hipMalloc(&d_reject, ns * sizeof(int)));
  struct Free { int* p; ~Free() { if (p) (void)hipFree(p); } } guard{d_reject};
  FZ_TRY(forces(true, true));
  fz_energies e;
  for (int s = 0; s < ns; s++) { sys_[s]->energies(&e); e_old[s] = e.potential; }
  for (int it = 0; it < nsteps; it++) {
    for (int s = 0; s < ns; s++)
      FZ_HIP(hipMemcpyAsync(sys_[s]->dv.mstep, &step[s], sizeof(double), hipMemcpyHostToDevice, L_.stream));
    launch_min_move(L_);
    FZ_HIP(hipGetLastError());
    FZ_TRY(forces(true, true));
    bool any = false;
    for (int s = 0; s < ns; s++) {
      sys_[s]->energies(&e);
      reject[s] = !(e.potential < e_old[s]);
      if (reject[s]) { step[s] *= 0.5; any = true; }
      else { e_old[s] = e.potential; step[s] = std::min(step[s] * 1.2, max_step); }
    }
    if (any) {
      FZ_HIP(hipMemcpyAsync(d_reject, reject.data(), ns * sizeof(int), hipMemcpyHostToDevice, L_.stream));
      launch_min_restore(L_, d_reject);
      FZ_TRY(forces(true, true));
    }
  }
  // the adaptive step carries over to the next call, so a minimisation split
  // into several calls takes the same path as one call
  for (int s = 0; s < ns; s++) sys_[s]->min_step = step[s];
  launch_zero_vel(L_);
  return forces(true);
}

void launch_min_ref_move(const Launch& L);

// The reference's schedule (MinimizeStep, src/dynamics.cu:491-620; factor ramp
// src/fenzi.cu:402-415), per system: move by the current method, evaluate; an
// energy above the last accepted one (the first move is always accepted) undoes
// the move and demotes conjugate gradient -> steepest descent -> stop; after
// every 100 accepted moves the active factor grows by rate * 100.
int Engine::minimize_cgsd(int nsteps, const fz_min_params& p) {
  const int ns = (int)sys_.size();
  std::vector<int> method(ns, p.method), done(ns, 0), have(ns, 0), reject(ns, 0);
  std::vector<double> cg(ns, p.cgfac), sd(ns, p.sdfac), e_prev(ns, 0.0);
  std::vector<double> ctl(2 * ns);
  int* d_reject = nullptr;
  FZ_HIP(hipMalloc(&d_reject, ns * sizeof(int)));
  struct Free { int* p; ~Free() { if (p) (void)hipFree(p); } } guard{d_reject};
  for (auto* s : sys_) {
    FZ_HIP(hipMemsetAsync(s->dv.mH, 0, sizeof(double) * 3 * s->topo.n, L_.stream));
    FZ_HIP(hipMemsetAsync(s->dv.mF, 0, sizeof(double) * 3 * s->topo.n, L_.stream));
  }
  FZ_TRY(forces(true, true));
  fz_energies e;
  for (int it = 0; it < nsteps; it++) {
    bool active = false;
    for (int s = 0; s < ns; s++) {
      active = active || method[s] != 0;
      ctl[2 * s] = method[s] == 1 ? cg[s] : sd[s];
      ctl[2 * s + 1] = (double)method[s];
      FZ_HIP(hipMemcpyAsync(sys_[s]->dv.mstep, &ctl[2 * s], 2 * sizeof(double), hipMemcpyHostToDevice, L_.stream));
    }
    if (!active) break;
    launch_min_ref_move(L_);
    FZ_HIP(hipGetLastError());
    FZ_TRY(forces(true, true));
    bool any = false;
    for (int s = 0; s < ns; s++) {
      reject[s] = 0;
      if (method[s] == 0) continue;
      sys_[s]->energies(&e);
      if (have[s] && e.potential > e_prev[s]) {
        reject[s] = 1;
        any = true;
        method[s] = method[s] == 1 ? 2 : 0;
      } else {
        e_prev[s] = e.potential;
        have[s] = 1;
        const int idx = done[s]++;  // 0-based index of the accepted move
        if (idx > 0 && idx % 100 == 0) {
          if (method[s] == 1) cg[s] *= p.cgrate * 100.0;
          else sd[s] *= p.sdrate * 100.0;
        }
      }
    }
    if (any) {
      FZ_HIP(hipMemcpyAsync(d_reject, reject.data(), ns * sizeof(int), hipMemcpyHostToDevice, L_.stream));
      launch_min_restore(L_, d_reject);
      FZ_TRY(forces(true, true));
    }
  }
  launch_zero_vel(L_);
  return forces(true);
}

// A failed minimisation leaves every member as it was before the call (a
// batch failure is then retried member by member from the same state).
int Engine::minimize(int nsteps, double max_step) {
  FZ_TRY(ensure());
  for (auto* s : sys_)
    if (s->need_rebuild) FZ_TRY(forces(true));
  FZ_TRY(snapshot());
  const int rc = minimize_sd(nsteps, max_step);
  if (rc != FZ_OK) {
    const std::string msg = last_error();
    (void)restore();
    return set_error(rc, "%s", msg.c_str());
  }
  return npt_reset_virial();
}

int Engine::minimize_ref(int nsteps, const fz_min_params& p) {
  if (p.method == 0) return minimize(nsteps, p.max_step);
  if (p.method != 1 && p.method != 2) return set_error(FZ_ERR_ARG, "unknown minimisation method %d", p.method);
  FZ_TRY(ensure());
  for (auto* s : sys_)
    if (s->need_rebuild) FZ_TRY(forces(true));
  FZ_TRY(snapshot());
  const int rc = minimize_cgsd(nsteps, p);
  if (rc != FZ_OK) {
    const std::string msg = last_error();
    (void)restore();
    return set_error(rc, "%s", msg.c_str());
  }
  return npt_reset_virial();
}

int Engine::pair_kernel_time(int reps, double* ms) {
  FZ_TRY(forces(true));  // a current list, slot copies and work items
  hipEvent_t e0, e1;
  FZ_HIP(hipEventCreate(&e0));
  FZ_HIP(hipEventCreate(&e1));
  (void)hipEventRecord(e0, L_.stream);
  for (int k = 0; k < reps; k++) launch_pair(L_, 0);
  (void)hipEventRecord(e1, L_.stream);
  const hipError_t rc = hipEventSynchronize(e1);
  float t = 0.f;
  if (rc == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != hipSuccess) return set_error(FZ_ERR_HIP, "event synchronize failed: %s", hipGetErrorString(rc));
  *ms = (double)t / reps;
  return forces(true);  // the repeated launches summed their forces several times over
}

int Engine::phase_times(int nsteps, double* out) {
  FZ_TRY(ensure());
  for (auto* s : sys_)
    if (s->need_rebuild) FZ_TRY(forces(true));
  FZ_TRY(snapshot(true));
  const int NP = 9;
  hipEvent_t ev[NP + 1];
  for (int k = 0; k <= NP; k++) FZ_HIP(hipEventCreate(&ev[k]));
  for (int k = 0; k < NP; k++) out[k] = 0.0;
  int rc = FZ_OK;
  for (int st = 0; st < nsteps && rc == FZ_OK; st++) {
    (void)hipEventRecord(ev[0], L_.stream);
    launch_step_first(L_);
    (void)hipEventRecord(ev[1], L_.stream);
    launch_nblist(L_, false);
    (void)hipEventRecord(ev[2], L_.stream);
    launch_gather_sorted(L_);
    (void)hipEventRecord(ev[3], L_.stream);
    launch_pair(L_, 0);
    (void)hipEventRecord(ev[4], L_.stream);
    launch_bonded(L_, false);
    (void)hipEventRecord(ev[5], L_.stream);
    launch_pme_part(L_, 0, false);
    (void)hipEventRecord(ev[6], L_.stream);
    launch_pme_part(L_, 1, false);
    (void)hipEventRecord(ev[7], L_.stream);
    launch_pme_part(L_, 2, false);
    (void)hipEventRecord(ev[8], L_.stream);
    launch_step_second(L_);
    (void)hipEventRecord(ev[9], L_.stream);
    if (hipEventSynchronize(ev[9]) != hipSuccess) rc = set_error(FZ_ERR_HIP, "event synchronize failed");
    for (int k = 0; k < NP && rc == FZ_OK; k++) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
      out[k] += ms;
    }
  }
  for (int k = 0; k <= NP; k++) (void)hipEventDestroy(ev[k]);
  FZ_TRY(rc);
  for (int k = 0; k < NP; k++) out[k] /= std::max(1, nsteps);
  FZ_TRY(read_status());
  for (size_t k = 0; k < sys_.size(); k++) {
    const int ov = sys_[k]->flags_host[F_OVERFLOW], sf = sys_[k]->flags_host[F_SHAKE_FAIL];
    if (ov || sf) {
      // eager steps do not replay: the members go back to where they were
      FZ_TRY(restore());
      return set_error(ov ? FZ_ERR_CAPACITY : FZ_ERR_SHAKE, "%s in system %d during the phase timing",
                       ov ? "neighbour capacity overflow" : "SHAKE failure", (int)k);
    }
  }
  for (auto* s : sys_) s->step += nsteps;
  return FZ_OK;
}

int Engine::time_steps(int nsteps, double* ms) {
  FZ_TRY(steps(2));  // graph capture and warm-up outside the timed region
  FZ_TRY(sync());
  hipEvent_t a, b;
  FZ_HIP(hipEventCreate(&a));
  FZ_HIP(hipEventCreate(&b));
  (void)hipEventRecord(a, L_.stream);
  int rc = steps(nsteps);
  (void)hipEventRecord(b, L_.stream);
  (void)hipEventSynchronize(b);
  float t = 0.f;
  (void)hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t / std::max(1, nsteps);
  return rc;
}

// Coordinate descent over the launch-shape knobs the reference fixes at
// compile time (CELL_BLOCKSIZE / CELL_ATOMS presets, src/configure:89-137):
// column width (atoms per column and cluster compactness), FFT lines per
// workgroup, persistent pair-kernel grid.  Each candidate is timed on
// graph-replayed steps from the same state; a candidate whose steps fail (a
// capacity or SHAKE error) counts as infinitely slow.  Members with an explicit
// column grid keep it.  On every exit the members' positions, velocities,
// step counters and minimiser steps are those they had on entry.
int Engine::set_launch(int nl, int grid, int pme_path, int32_t* out) {
  if (nl > 64 || grid > 65536 || pme_path > FZ_PME_PLANES) return set_error(FZ_ERR_ARG, "invalid launch shape");
  if (nl >= 0) tuned_nl_ = nl;
  if (grid >= 0) tuned_grid_ = grid;
  if (pme_path >= 0) tuned_planes_ = pme_path == FZ_PME_AUTO ? -1 : pme_path == FZ_PME_PLANES ? 1 : 0;
  FZ_TRY(prepare());
  if (tuned_planes_ == 1 && !L_.pme_planes) {
    tuned_planes_ = -1;
    FZ_TRY(prepare());
    return set_error(FZ_ERR_ARG, "a PME grid plane of this set does not fit in %d bytes of LDS", PLANE_MAX_LDS);
  }
  if (out) {
    out[0] = L_.fft_nl;
    out[1] = L_.pair_grid;
    out[2] = L_.pme_planes ? FZ_PME_PLANES : FZ_PME_GRID;
  }
  return FZ_OK;
}

int Engine::autotune(int nsteps, double* out) {
  FZ_TRY(ensure());
  const int ns = (int)sys_.size();
  std::vector<std::vector<uint64_t>> u0(ns);
  std::vector<std::vector<double>> v0(ns);
  std::vector<int64_t> st0(ns);
  std::vector<double> ms0(ns);
  for (int s = 0; s < ns; s++) {
    u0[s].resize((size_t)3 * sys_[s]->topo.n);
    v0[s].resize((size_t)3 * sys_[s]->topo.n);
    FZ_TRY(sys_[s]->download_state(nullptr, v0[s].data(), u0[s].data()));
    st0[s] = sys_[s]->step;
    ms0[s] = sys_[s]->min_step;
  }
  auto restore_state = [&]() -> int {
    for (int s = 0; s < ns; s++) {
      FZ_TRY(sys_[s]->upload_fixed(u0[s].data()));
      FZ_TRY(sys_[s]->upload_velocities(v0[s].data()));
      sys_[s]->step = st0[s];
      sys_[s]->min_step = ms0[s];
    }
    return FZ_OK;
  };
  auto configure = [&](double f, int nl, int grid) -> int {
    for (auto* s : sys_) FZ_TRY(s->set_column_grid(s->par.column_grid[0], s->par.column_grid[1], f));
    tuned_nl_ = nl;
    tuned_grid_ = grid;
    FZ_TRY(prepare());
    return restore_state();
  };
  auto trial = [&](double f, int nl, int grid, double* ms) -> int {
    FZ_TRY(configure(f, nl, grid));
    const int rc = time_steps(nsteps, ms);
    if (rc == FZ_ERR_SHAKE || rc == FZ_ERR_CAPACITY) *ms = 1e30;
    else if (rc != FZ_OK) return rc;
    return FZ_OK;
  };
  bool explicit_grid = true;
  for (auto* s : sys_) explicit_grid = explicit_grid && s->par.column_grid[0] > 0 && s->par.column_grid[1] > 0;
  double best_f = 1.0, best_ms = 1e30, ms = 0;
  int best_nl = 16, best_grid = 2048;
  int rc = FZ_OK;
  std::vector<double> widths = {0.8, 0.9, 1.0, 1.15, 1.3};
  if (explicit_grid) widths = {1.0};
  for (double f : widths) {
    if ((rc = trial(f, best_nl, best_grid, &ms)) != FZ_OK) break;
    if (ms < best_ms) { best_ms = ms; best_f = f; }
  }
  for (int nl : {4, 8, 16}) {
    if (rc != FZ_OK || (rc = trial(best_f, nl, best_grid, &ms)) != FZ_OK) break;
    if (ms < best_ms) { best_ms = ms; best_nl = nl; }
  }
  for (int grid : {512, 1024, 4096, 8192, 16384, 32768}) {
    if (rc != FZ_OK || (rc = trial(best_f, best_nl, grid, &ms)) != FZ_OK) break;
    if (ms < best_ms) { best_ms = ms; best_grid = grid; }
  }
  if (rc != FZ_OK) {
    // an unexpected error: back to the defaults and the entry state
    const std::string msg = last_error();
    tuned_nl_ = tuned_grid_ = 0;
    (void)configure(1.0, 0, 0);
    (void)forces(true);
    return set_error(rc, "%s", msg.c_str());
  }
  FZ_TRY(configure(best_f, best_nl, best_grid));
  FZ_TRY(forces(true));
  out[0] = best_f; out[1] = best_nl; out[2] = best_grid; out[3] = best_ms;
  return FZ_OK;
}

int Engine::maxwell(double T, uint64_t seed, uint64_t counter) {
  FZ_TRY(ensure());
  launch_maxwell(L_, T, seed, counter);
  launch_zero_acc(L_);
  launch_kinetic(L_);
  FZ_HIP(hipGetLastError());
  FZ_TRY(read_status());
  return FZ_OK;
}

// ---------------------------------------------------------------------------
// stand-alone PME convolution (test entry)

int pme_convolve(int device, const int32_t K[3], const double box[3], double beta, int mixed, int mode,
                 const double* q, double* phi, double* energy) {
  FZ_HIP(hipSetDevice(device));
  for (int d = 0; d < 3; d++)
    if (!smooth235(K[d]) || K[d] > MAX_FFT || K[d] < 2) return set_error(FZ_ERR_ARG, "grid must be 2,3,5-smooth, 2..%d", MAX_FFT);
  System s;
  s.device = device;
  s.mixed = mixed;
  SysDev& d = s.dv;
  memset(&d, 0, sizeof d);
  for (int k = 0; k < 3; k++) { d.K[k] = K[k]; d.L[k] = box[k]; }
  d.Kxh = K[0] / 2 + 1;
  d.beta = beta;
  const double V = box[0] * box[1] * box[2];
  d.pme_pre = kCC / (kPi * V);
  d.pi2_beta2 = kPi * kPi / (beta * beta);
  const size_t K3 = (size_t)K[0] * K[1] * K[2], KH = (size_t)d.Kxh * K[1] * K[2];
  const int KS = K[0] + K[1] + K[2];
  void* p;
  FZ_TRY(s.dalloc((unsigned long long**)&p, K3)); d.Qacc = p;
  FZ_TRY(s.dalloc((char**)&p, KH * (mixed ? 8 : 16))); d.Qc = p;
  FZ_TRY(s.dalloc((char**)&p, K3 * (mixed ? 4 : 8))); d.phi = p;
  std::vector<double> bsp(KS), ex(KS), mv(KS);
  std::vector<double2> tw64(KS);
  std::vector<float2> tw32(KS);
  int off = 0;
  for (int ax = 0; ax < 3; ax++) {
    for (int j = 0; j < K[ax]; j++) {
      const int jj = j <= K[ax] / 2 ? j : j - K[ax];
      const double m = jj / box[ax], c = std::cos(2.0 * kPi * j / K[ax]);
      bsp[off + j] = 9.0 / ((2.0 + c) * (2.0 + c));
      ex[off + j] = std::exp(-kPi * kPi * m * m / (beta * beta));
      mv[off + j] = m;
      const double a = -2.0 * kPi * j / K[ax];
      tw64[off + j] = make_double2(std::cos(a), std::sin(a));
      tw32[off + j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    off += K[ax];
  }
  FZ_TRY(s.dalloc(&d.pme_bsp, KS)); FZ_TRY(up(d.pme_bsp, bsp));
  FZ_TRY(s.dalloc(&d.pme_ex, KS)); FZ_TRY(up(d.pme_ex, ex));
  FZ_TRY(s.dalloc(&d.pme_m, KS)); FZ_TRY(up(d.pme_m, mv));
  if (mixed) { float2* t; FZ_TRY(s.dalloc(&t, KS)); FZ_TRY(up(t, tw32)); d.tw = t; }
  else { double2* t; FZ_TRY(s.dalloc(&t, KS)); FZ_TRY(up(t, tw64)); d.tw = t; }
  FZ_TRY(s.dalloc(&d.acc, A_NSLOT));
  FZ_TRY(s.dalloc(&d.flags, F_NSLOT));
  if (mixed) {
    std::vector<unsigned long long> h(K3);
    for (size_t i = 0; i < K3; i++) h[i] = (unsigned long long)llrint(q[i] * 4294967296.0);
    FZ_HIP(hipMemcpy(d.Qacc, h.data(), K3 * 8, hipMemcpyHostToDevice));
  } else {
    FZ_HIP(hipMemcpy(d.Qacc, q, K3 * 8, hipMemcpyHostToDevice));
  }
  Engine eng(device, {&s});
  FZ_TRY(eng.prepare());
  launch_pme_fft_only(eng.launch(), mode);
  FZ_HIP(hipGetLastError());
  FZ_TRY(eng.read_status());
  if (mixed) {
    std::vector<float> h(K3);
    FZ_HIP(hipMemcpy(h.data(), d.phi, K3 * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < K3; i++) phi[i] = h[i];
  } else {
    FZ_HIP(hipMemcpy(phi, d.phi, K3 * 8, hipMemcpyDeviceToHost));
  }
  if (energy) *energy = fx(s.acc_host[A_PME]);
  return FZ_OK;
}

}  // namespace fz
