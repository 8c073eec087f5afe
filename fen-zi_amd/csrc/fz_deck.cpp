// fz_deck.cpp -- FEN ZI input surface: .in keyword file, CHARMM .psf / .prm /
// .rtf, .crd / .pdb / .xyz coordinates, and the energy-table output format.
//
// Replaces ReadInput (src/prmread.cu:491-913), ReadCrd / ReadPdb / ReadXYZ
// (:331, :410, :285), ReadPsf_CHARMM (:1442-1955), ReadPrm_CHARMM
// (:2181-2750), build_constraints (:1192-1303) and print_output_headers /
// EvalProps' row printf (src/evalprops.cu:248-267, 311-343).  Kept: keyword
// names and defaults (src/globals.cu), case-insensitive keys, '#' comments,
// numeric PSF types mapped through the RTF MASS table, per-type masses taken
// from the first atom of each type (src/prmread.cu:1580-1602), PRM unit
// conversions (theta0 and delta to radians; sigma = Rmin/2 / 2^(1/6);
// NBFIX eps = |Emin|, sigma = Rmin / 2^(1/6)), SHAKE constraints on bonds with a
// light (<= 1.5 amu) atom.  Changed on purpose: PSF sections are located by
// their !N... headers (the reference reads molecule ids from the wrong
// section after NIMPHI), the energy bonded-parameter tables are stored per term
// (no per-type hash caps), and dihedral wildcards follow the CHARMM X-B-C-X
// rule with specific entries taking precedence.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fenzi.h"

namespace fz {
int set_error(int code, const char* fmt, ...);
}
using fz::set_error;

namespace {

const double kDeg = 3.14159265358979323846 / 180.0;

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}
std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

std::vector<std::string> split(const std::string& line) {
  std::vector<std::string> t;
  std::istringstream is(line);
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

std::string strip_comment(const std::string& line, char c) {
  size_t p = line.find(c);
  return p == std::string::npos ? line : line.substr(0, p);
}

std::string dir_of(const std::string& path) {
  size_t p = path.find_last_of('/');
  return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

std::string resolve(const std::string& base, const std::string& name) {
  if (name.empty() || name[0] == '/') return name;
  std::ifstream f(name);
  if (f.good()) return name;
  return base + name;
}

struct BondPrm { double k, r0; };
struct AnglePrm { double k, t0, kub, s0; };
struct DihTerm { double k; int n; double delta; };

}  // namespace

struct fz_deck {
  fz_topology topo{};
  fz_params par{};
  fz_run_settings set{};
  std::vector<double> xyz;
  // owned topology arrays
  std::vector<double> charge, mass, lj_eps, lj_sig2, lj14_eps, lj14_sig2;
  std::vector<int32_t> type;
  std::vector<int32_t> bond_atoms, angle_atoms, dihedral_atoms, dihedral_n, improper_atoms, constraint_atoms;
  std::vector<double> bond_k, bond_r0, angle_k, angle_theta0, ub_k, ub_s0, dihedral_k, dihedral_delta,
      improper_k, improper_psi0, constraint_length;
  std::vector<std::string> type_names;
  int64_t psf_counts[4] = {0, 0, 0, 0};
  // segment names: CRD column after the coordinates (the consharm groups,
  // src/prmread.cu:366-378) and PSF column 2 (the consfix segments, :1549)
  std::vector<std::string> crd_seg, psf_seg;
  bool consharm = false;
  std::string consharm_seg[2];
  double consharm_dist[3] = {10.0, 10.0, 20.0};  // src/globals.cu:231-232
  double consharm_fc = 0.01;
  std::vector<std::string> consfix_seg;
};

namespace {

struct Prm {
  std::map<std::string, BondPrm> bonds;
  std::map<std::string, AnglePrm> angles;
  std::map<std::string, std::vector<DihTerm>> diheds;
  std::map<std::string, std::pair<double, double>> imprs;  // k, psi0
  std::map<std::string, double> nb_eps, nb_rmin2, nb14_eps, nb14_rmin2;
  std::map<std::string, std::pair<double, double>> nbfix;  // eps (>0), rmin
};

std::string key2(const std::string& a, const std::string& b) { return a + " " + b; }
std::string key3(const std::string& a, const std::string& b, const std::string& c) { return a + " " + b + " " + c; }
std::string key4(const std::string& a, const std::string& b, const std::string& c, const std::string& d) {
  return a + " " + b + " " + c + " " + d;
}

int read_prm(const std::string& path, Prm& P) {
  std::ifstream f(path);
  if (!f) return set_error(FZ_ERR_PARSE, "parameter file %s not found", path.c_str());
  std::string line, sec;
  std::string cont;
  std::map<std::string, bool> dih_seen_in_block;
  while (std::getline(f, line)) {
    std::string s = strip_comment(line, '!');
    auto t = split(s);
    if (t.empty()) continue;
    std::string h = upper(t[0]);
    if (h == "BONDS" || h == "BOND") { sec = "B"; continue; }
    if (h == "ANGLES" || h == "ANGL" || h == "THETAS") { sec = "A"; continue; }
    if (h == "DIHEDRALS" || h == "DIHE" || h == "PHI") { sec = "D"; continue; }
    if (h == "IMPROPER" || h == "IMPROPERS" || h == "IMPR" || h == "IMPHI") { sec = "I"; continue; }
    if (h == "CMAP") { sec = "C"; continue; }
    if (h == "NONBONDED" || h == "NONB" || h == "NBONDED") {
      sec = "N";
      // header options may continue on following lines ending with '-'
      std::string cur = s;
      while (!cur.empty() && cur.find_last_not_of(" \t\r") != std::string::npos &&
             cur[cur.find_last_not_of(" \t\r")] == '-' && std::getline(f, line))
        cur = strip_comment(line, '!');
      continue;
    }
    if (h == "NBFIX") { sec = "F"; continue; }
    if (h == "HBOND") { sec = "H"; continue; }
    if (h == "END") break;
    if (sec == "B" && t.size() >= 4) {
      P.bonds[key2(t[0], t[1])] = BondPrm{atof(t[2].c_str()), atof(t[3].c_str())};
    } else if (sec == "A" && t.size() >= 5) {
      AnglePrm a{atof(t[3].c_str()), atof(t[4].c_str()) * kDeg, 0.0, 0.0};
      if (t.size() >= 7) { a.kub = atof(t[5].c_str()); a.s0 = atof(t[6].c_str()); }
      P.angles[key3(t[0], t[1], t[2])] = a;
    } else if (sec == "D" && t.size() >= 7) {
      std::string k = key4(t[0], t[1], t[2], t[3]);
      DihTerm d{atof(t[4].c_str()), atoi(t[5].c_str()), atof(t[6].c_str()) * kDeg};
      auto& v = P.diheds[k];
      // a repeated quadruple adds a term; a repeated multiplicity replaces it
      bool done = false;
      for (auto& e : v)
        if (e.n == d.n) { e = d; done = true; }
      if (!done) v.push_back(d);
    } else if (sec == "I" && t.size() >= 7) {
      P.imprs[key4(t[0], t[1], t[2], t[3])] = {atof(t[4].c_str()), atof(t[6].c_str()) * kDeg};
    } else if (sec == "N" && t.size() >= 4) {
      const std::string& a = t[0];
      double eps = std::fabs(atof(t[2].c_str())), r2 = atof(t[3].c_str());
      P.nb_eps[a] = eps;
      P.nb_rmin2[a] = r2;
      if (t.size() >= 7) {
        P.nb14_eps[a] = std::fabs(atof(t[5].c_str()));
        P.nb14_rmin2[a] = atof(t[6].c_str());
      } else {
        P.nb14_eps[a] = eps;
        P.nb14_rmin2[a] = r2;
      }
    } else if (sec == "F" && t.size() >= 4) {
      double e = std::fabs(atof(t[2].c_str())), r = atof(t[3].c_str());
      P.nbfix[key2(t[0], t[1])] = {e, r};
      P.nbfix[key2(t[1], t[0])] = {e, r};
    }
  }
  return FZ_OK;
}

int read_rtf_types(const std::string& path, std::map<int, std::string>& types) {
  std::ifstream f(path);
  if (!f) return FZ_ERR_PARSE;
  std::string line;
  while (std::getline(f, line)) {
    auto t = split(strip_comment(line, '!'));
    if (t.size() >= 3 && upper(t[0]) == "MASS") types[atoi(t[1].c_str())] = t[2];
  }
  return FZ_OK;
}

int read_crd(const std::string& path, std::vector<double>& xyz, std::vector<std::string>* seg = nullptr) {
  std::ifstream f(path);
  if (!f) return set_error(FZ_ERR_PARSE, "coordinate file %s not found", path.c_str());
  std::string line;
  long n = -1;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '*') continue;
    auto t = split(line);
    if (t.empty()) continue;
    n = atol(t[0].c_str());
    break;
  }
  if (n <= 0) return set_error(FZ_ERR_PARSE, "%s: missing atom count", path.c_str());
  xyz.resize(3 * n);
  for (long i = 0; i < n; i++) {
    if (!std::getline(f, line)) return set_error(FZ_ERR_PARSE, "%s: expected %ld atoms, found %ld", path.c_str(), n, i);
    auto t = split(line);
    if (t.size() < 7) return set_error(FZ_ERR_PARSE, "%s: malformed line %ld", path.c_str(), i + 1);
    for (int k = 0; k < 3; k++) xyz[3 * i + k] = atof(t[4 + k].c_str());
    if (seg) seg->push_back(t.size() > 7 ? t[7] : std::string());
  }
  return FZ_OK;
}

int read_pdb(const std::string& path, std::vector<double>& xyz) {
  std::ifstream f(path);
  if (!f) return set_error(FZ_ERR_PARSE, "coordinate file %s not found", path.c_str());
  std::string line;
  xyz.clear();
  while (std::getline(f, line)) {
    if (line.compare(0, 4, "ATOM") != 0 && line.compare(0, 6, "HETATM") != 0) continue;
    if (line.size() < 54) return set_error(FZ_ERR_PARSE, "%s: short ATOM record", path.c_str());
    xyz.push_back(atof(line.substr(30, 8).c_str()));
    xyz.push_back(atof(line.substr(38, 8).c_str()));
    xyz.push_back(atof(line.substr(46, 8).c_str()));
  }
  return FZ_OK;
}

int read_xyz(const std::string& path, std::vector<double>& xyz) {
  std::ifstream f(path);
  if (!f) return set_error(FZ_ERR_PARSE, "coordinate file %s not found", path.c_str());
  std::string line;
  std::getline(f, line);
  long n = atol(line.c_str());
  std::getline(f, line);
  xyz.resize(3 * n);
  for (long i = 0; i < n; i++) {
    if (!std::getline(f, line)) return set_error(FZ_ERR_PARSE, "%s: truncated", path.c_str());
    auto t = split(line);
    if (t.size() < 4) return set_error(FZ_ERR_PARSE, "%s: malformed line", path.c_str());
    for (int k = 0; k < 3; k++) xyz[3 * i + k] = atof(t[1 + k].c_str());
  }
  return FZ_OK;
}

bool find_dihedral(const Prm& P, const std::string& a, const std::string& b, const std::string& c,
                   const std::string& d, const std::vector<DihTerm>** out) {
  for (const auto& k : {key4(a, b, c, d), key4(d, c, b, a), key4("X", b, c, "X"), key4("X", c, b, "X")}) {
    auto it = P.diheds.find(k);
    if (it != P.diheds.end()) { *out = &it->second; return true; }
  }
  return false;
}

bool find_improper(const Prm& P, const std::string& a, const std::string& b, const std::string& c,
                   const std::string& d, std::pair<double, double>* out) {
  for (const auto& k : {key4(a, b, c, d), key4(d, c, b, a), key4(a, "X", "X", d), key4(d, "X", "X", a),
                        key4("X", b, c, d), key4("X", c, b, a), key4("X", "X", c, d), key4("X", "X", b, a)}) {
    auto it = P.imprs.find(k);
    if (it != P.imprs.end()) { *out = it->second; return true; }
  }
  return false;
}

// PSF: sections located by their "!N..." headers
int read_psf(const std::string& path, const std::map<int, std::string>* rtf, fz_deck* D,
             std::vector<std::string>& atom_type_name, std::vector<std::array<int, 4>>* secs[4]) {
  (void)secs;
  std::ifstream f(path);
  if (!f) return set_error(FZ_ERR_PARSE, "structure file %s not found", path.c_str());
  std::string line;
  std::vector<std::string> tok;
  // read all tokens after locating sections
  std::vector<std::string> lines;
  while (std::getline(f, line)) lines.push_back(line);
  size_t li = 0;
  auto find_sec = [&](const char* tag, long* count) -> bool {
    for (; li < lines.size(); li++) {
      size_t p = lines[li].find(tag);
      if (p != std::string::npos) {
        *count = atol(lines[li].c_str());
        li++;
        return true;
      }
    }
    return false;
  };
  long natom = 0;
  if (!find_sec("!NATOM", &natom) || natom <= 0) return set_error(FZ_ERR_PARSE, "%s: no !NATOM section", path.c_str());
  atom_type_name.resize(natom);
  D->charge.resize(natom);
  D->mass.resize(natom);
  for (long i = 0; i < natom; i++, li++) {
    if (li >= lines.size()) return set_error(FZ_ERR_PARSE, "%s: truncated atom section", path.c_str());
    auto t = split(lines[li]);
    if (t.size() < 8) return set_error(FZ_ERR_PARSE, "%s: malformed atom line %ld", path.c_str(), i + 1);
    std::string ty = t[5];
    if (rtf) {
      auto it = rtf->find(atoi(ty.c_str()));
      if (it == rtf->end()) return set_error(FZ_ERR_PARSE, "%s: type id %s not in topology file", path.c_str(), ty.c_str());
      ty = it->second;
    }
    atom_type_name[i] = ty;
    D->psf_seg.push_back(t[1]);
    D->charge[i] = atof(t[6].c_str());
    D->mass[i] = atof(t[7].c_str());
  }
  const char* tags[4] = {"!NBOND", "!NTHETA", "!NPHI", "!NIMPHI"};
  const int width[4] = {2, 3, 4, 4};
  std::vector<int32_t>* dst[4] = {&D->bond_atoms, &D->angle_atoms, &D->dihedral_atoms, &D->improper_atoms};
  std::vector<std::vector<int32_t>> raw(4);
  for (int s = 0; s < 4; s++) {
    long cnt = 0;
    size_t save = li;
    if (!find_sec(tags[s], &cnt)) {
      li = save;
      continue;
    }
    D->psf_counts[s] = cnt;
    long need = cnt * width[s];
    raw[s].reserve(need);
    while ((long)raw[s].size() < need && li < lines.size()) {
      for (auto& w : split(lines[li])) raw[s].push_back(atoi(w.c_str()) - 1);
      li++;
    }
    if ((long)raw[s].size() < need) return set_error(FZ_ERR_PARSE, "%s: truncated %s section", path.c_str(), tags[s]);
    raw[s].resize(need);
    for (auto a : raw[s])
      if (a < 0 || a >= natom) return set_error(FZ_ERR_PARSE, "%s: atom index out of range in %s", path.c_str(), tags[s]);
    *dst[s] = raw[s];
  }
  return FZ_OK;
}

}  // namespace

extern "C" {

int fz_deck_load(const char* in_path, fz_deck** out) {
  try {
    if (!in_path || !out) return set_error(FZ_ERR_ARG, "null argument");
    std::ifstream f(in_path);
    if (!f) return set_error(FZ_ERR_PARSE, "input file %s not found", in_path);
    auto* D = new fz_deck();
    std::unique_ptr<fz_deck> guard(D);
    const std::string base = dir_of(in_path);
    // defaults: src/globals.cu
    double T = 298.0, dt = 0.001, kappa = 0.032, cuton = 7.0, cutoff = 8.0, pairlist = 9.5, shaketol = 1e-6;
    long dynasteps = 10000, outputTiming = 100, startstep = 0, minimize = 0, tempreset = 0;
    int shake = 0, printstdout = 1;
    double seed = 0;
    double cgfac = 1e-10, sdfac = 1e-4, cgrate = 0.1, sdrate = 0.01;
    long ckpt_every = 100000, traj_every = -1;
    int printlevel = 0, ckpt_binary = 1;
    double pmass = 0.0, pref = 1.0;  // barostat off unless pmass is given (PREF 1 atm, src/defs.in.h:101)
    double box[3] = {0, 0, 0};
    int K[3] = {0, 0, 0};
    std::string crd, psf, prm, rtf, outname;
    std::string line;
    while (std::getline(f, line)) {
      auto t = split(line);
      if (t.empty() || t[0][0] == '#') continue;
      std::string k = lower(t[0]);
      auto val = [&](size_t i) -> std::string { return i < t.size() ? t[i] : std::string(); };
      if (k == "coordinates") crd = val(1);
      else if (k == "structure") psf = val(1);
      else if (k == "parameters") prm = val(1);
      else if (k == "topology") rtf = val(1);
      else if (k == "temperature") T = atof(val(1).c_str());
      else if (k == "minimizesteps") minimize = (long)atof(val(1).c_str());
      else if (k == "conjgradfac") cgfac = atof(val(1).c_str());
      else if (k == "steepdesfac") sdfac = atof(val(1).c_str());
      else if (k == "conjgradrate") cgrate = atof(val(1).c_str());
      else if (k == "steepdesrate") sdrate = atof(val(1).c_str());
      else if (k == "startstep") startstep = atol(val(1).c_str());
      else if (k == "timestep") dt = atof(val(1).c_str());
      else if (k == "dynasteps") dynasteps = atol(val(1).c_str());
      else if (k == "kappa") kappa = atof(val(1).c_str());
      else if (k == "tempresettiming") tempreset = atol(val(1).c_str());
      else if (k == "outputtiming") outputTiming = atol(val(1).c_str());
      else if (k == "cuton") cuton = atof(val(1).c_str());
      else if (k == "cutoff") cutoff = atof(val(1).c_str());
      else if (k == "pairlistdist") pairlist = atof(val(1).c_str());
      else if (k == "cellbasisvector1") box[0] = atof(val(1).c_str());
      else if (k == "cellbasisvector2") box[1] = atof(val(2).c_str());
      else if (k == "cellbasisvector3") box[2] = atof(val(3).c_str());
      else if (k == "pmegridsizex") K[0] = atoi(val(1).c_str());
      else if (k == "pmegridsizey") K[1] = atoi(val(1).c_str());
      else if (k == "pmegridsizez") K[2] = atoi(val(1).c_str());
      else if (k == "shake") {
        shake = lower(val(1)) == "on";
        if (!val(2).empty()) shaketol = atof(val(2).c_str());
      } else if (k == "seed") seed = atof(val(1).c_str());
      else if (k == "printtostdout") printstdout = atoi(val(1).c_str());
      else if (k == "outputtrajectory" || k == "outputname") { if (outname.empty() || k == "outputtrajectory") outname = val(1); }
      else if (k == "checkpointtiming") ckpt_every = atol(val(1).c_str());
      else if (k == "trajfrequencytiming") traj_every = atol(val(1).c_str());
      else if (k == "printlevel") printlevel = atoi(val(1).c_str());
      else if (k == "checkpointformat") {
        const char c = val(1).empty() ? 'b' : val(1)[0];
        if (c == 'a' || c == 'A' || c == '1') ckpt_binary = 0;
        else if (c == 'b' || c == 'B' || c == '2') ckpt_binary = 1;
        else return set_error(FZ_ERR_PARSE, "checkpoint format \"%s\" is undefined (Ascii or Binary)", val(1).c_str());
      }
      // restraint and fixed segments (src/prmread.cu:674-690, 699-722)
      else if (k == "consharm") {
        D->consharm = true;
        D->consharm_seg[0] = val(1);
        D->consharm_seg[1] = val(2);
      } else if (k == "consharmdist") {
        for (int c = 0; c < 3; c++) D->consharm_dist[c] = atof(val(1 + c).c_str());
      } else if (k == "consharmfc") {
        D->consharm_fc = atof(val(1).c_str());
      } else if (k == "consfix") {
        const int ns = atoi(val(1).c_str());
        for (int j = 0; j < ns; j++) D->consfix_seg.push_back(val(2 + j));
      }
      // Andersen barostat (src/prmread.cu:664-670)
      else if (k == "pmass") pmass = atof(val(1).c_str());
      else if (k == "pref") pref = atof(val(1).c_str());
      // accepted and not used by the force-and-integrate path
      else if (k == "thermostat" || k == "inbfrq" || k == "gpublocksize" || k == "hfac" ||
               k == "keepcheckpoints" || k == "restartfilename") {
      } else {
        fprintf(stderr, "Ignoring unrecognized token, \"%s\"\n", t[0].c_str());
      }
    }
    if (crd.empty() || psf.empty() || prm.empty()) return set_error(FZ_ERR_PARSE, "%s: coordinates, structure and parameters are required", in_path);
    for (int d = 0; d < 3; d++)
      if (!(box[d] > 0)) return set_error(FZ_ERR_PARSE, "%s: cellBasisVector%d missing", in_path, d + 1);
    // coordinates
    std::string cpath = resolve(base, crd);
    std::string ext = lower(cpath.substr(cpath.find_last_of('.') + 1));
    int r = ext == "pdb" ? read_pdb(cpath, D->xyz) : ext == "xyz" ? read_xyz(cpath, D->xyz) : read_crd(cpath, D->xyz, &D->crd_seg);
    if (r != FZ_OK) return r;
    // topology type names
    std::map<int, std::string> rtf_types;
    bool have_rtf = false;
    if (!rtf.empty() && read_rtf_types(resolve(base, rtf), rtf_types) == FZ_OK) have_rtf = true;
    std::vector<std::string> tname;
    std::vector<std::array<int, 4>>* secs[4] = {nullptr, nullptr, nullptr, nullptr};
    r = read_psf(resolve(base, psf), have_rtf ? &rtf_types : nullptr, D, tname, secs);
    if (r != FZ_OK) return r;
    const int n = (int)D->charge.size();
    if ((long)D->xyz.size() != 3L * n)
      return set_error(FZ_ERR_PARSE, "Number of atoms in psf (%d) does not match number of atoms in coord file (%ld)", n,
                       (long)D->xyz.size() / 3);
    Prm P;
    r = read_prm(resolve(base, prm), P);
    if (r != FZ_OK) return r;
    // types in order of appearance; per-type mass from the first atom
    std::map<std::string, int> tid;
    std::vector<double> tmass;
    D->type.resize(n);
    for (int i = 0; i < n; i++) {
      auto it = tid.find(tname[i]);
      if (it == tid.end()) {
        it = tid.emplace(tname[i], (int)D->type_names.size()).first;
        D->type_names.push_back(tname[i]);
        tmass.push_back(D->mass[i]);
      }
      D->type[i] = it->second;
      D->mass[i] = tmass[it->second];
    }
    const int T_ = (int)D->type_names.size();
    D->lj_eps.assign(T_ * T_, 0.0); D->lj_sig2.assign(T_ * T_, 0.0);
    D->lj14_eps.assign(T_ * T_, 0.0); D->lj14_sig2.assign(T_ * T_, 0.0);
    const double s6 = std::pow(2.0, 1.0 / 6.0);
    for (int a = 0; a < T_; a++) {
      const std::string& na = D->type_names[a];
      if (!P.nb_eps.count(na)) return set_error(FZ_ERR_PARSE, "no NONBONDED parameters for type %s", na.c_str());
      for (int b = 0; b < T_; b++) {
        const std::string& nb = D->type_names[b];
        if (!P.nb_eps.count(nb)) return set_error(FZ_ERR_PARSE, "no NONBONDED parameters for type %s", nb.c_str());
        double sig = P.nb_rmin2[na] / s6 + P.nb_rmin2[nb] / s6;
        double eps = 4.0 * std::sqrt(P.nb_eps[na] * P.nb_eps[nb]);
        auto fx = P.nbfix.find(key2(na, nb));
        if (fx != P.nbfix.end()) { eps = 4.0 * fx->second.first; sig = fx->second.second / s6; }
        D->lj_eps[a * T_ + b] = eps;
        D->lj_sig2[a * T_ + b] = sig * sig;
        double sig14 = P.nb14_rmin2[na] / s6 + P.nb14_rmin2[nb] / s6;
        D->lj14_eps[a * T_ + b] = 4.0 * std::sqrt(P.nb14_eps[na] * P.nb14_eps[nb]);
        D->lj14_sig2[a * T_ + b] = sig14 * sig14;
      }
    }
    auto tn = [&](int atom) -> const std::string& { return D->type_names[D->type[atom]]; };
    // bonds
    const int nb = (int)D->bond_atoms.size() / 2;
    for (int i = 0; i < nb; i++) {
      int a = D->bond_atoms[2 * i], b = D->bond_atoms[2 * i + 1];
      auto it = P.bonds.find(key2(tn(a), tn(b)));
      if (it == P.bonds.end()) it = P.bonds.find(key2(tn(b), tn(a)));
      if (it == P.bonds.end()) return set_error(FZ_ERR_PARSE, "no bond parameters for %s-%s", tn(a).c_str(), tn(b).c_str());
      D->bond_k.push_back(it->second.k);
      D->bond_r0.push_back(it->second.r0);
      if (shake && (D->mass[a] <= 1.5 || D->mass[b] <= 1.5)) {
        D->constraint_atoms.push_back(a);
        D->constraint_atoms.push_back(b);
        D->constraint_length.push_back(it->second.r0);
      }
    }
    const int na = (int)D->angle_atoms.size() / 3;
    for (int i = 0; i < na; i++) {
      int a = D->angle_atoms[3 * i], b = D->angle_atoms[3 * i + 1], c = D->angle_atoms[3 * i + 2];
      auto it = P.angles.find(key3(tn(a), tn(b), tn(c)));
      if (it == P.angles.end()) it = P.angles.find(key3(tn(c), tn(b), tn(a)));
      if (it == P.angles.end())
        return set_error(FZ_ERR_PARSE, "no angle parameters for %s-%s-%s", tn(a).c_str(), tn(b).c_str(), tn(c).c_str());
      D->angle_k.push_back(it->second.k);
      D->angle_theta0.push_back(it->second.t0);
      D->ub_k.push_back(it->second.kub);
      D->ub_s0.push_back(it->second.s0);
    }
    std::vector<int32_t> dih_raw = D->dihedral_atoms;
    D->dihedral_atoms.clear();
    for (size_t i = 0; i < dih_raw.size() / 4; i++) {
      const int32_t* q = &dih_raw[4 * i];
      const std::vector<DihTerm>* terms = nullptr;
      if (!find_dihedral(P, tn(q[0]), tn(q[1]), tn(q[2]), tn(q[3]), &terms))
        return set_error(FZ_ERR_PARSE, "no dihedral parameters for %s-%s-%s-%s", tn(q[0]).c_str(), tn(q[1]).c_str(),
                         tn(q[2]).c_str(), tn(q[3]).c_str());
      for (const auto& tm : *terms) {
        D->dihedral_atoms.insert(D->dihedral_atoms.end(), q, q + 4);
        D->dihedral_k.push_back(tm.k);
        D->dihedral_n.push_back(tm.n);
        D->dihedral_delta.push_back(tm.delta);
      }
    }
    const int ni = (int)D->improper_atoms.size() / 4;
    for (int i = 0; i < ni; i++) {
      const int32_t* q = &D->improper_atoms[4 * i];
      std::pair<double, double> p;
      if (!find_improper(P, tn(q[0]), tn(q[1]), tn(q[2]), tn(q[3]), &p))
        return set_error(FZ_ERR_PARSE, "no improper parameters for %s-%s-%s-%s", tn(q[0]).c_str(), tn(q[1]).c_str(),
                         tn(q[2]).c_str(), tn(q[3]).c_str());
      D->improper_k.push_back(p.first);
      D->improper_psi0.push_back(p.second);
    }
    fz_topology& t = D->topo;
    t.n_atoms = n;
    t.n_types = T_;
    t.charge = D->charge.data(); t.mass = D->mass.data(); t.type = D->type.data();
    t.lj_eps = D->lj_eps.data(); t.lj_sig2 = D->lj_sig2.data();
    t.lj14_eps = D->lj14_eps.data(); t.lj14_sig2 = D->lj14_sig2.data();
    t.n_bonds = nb; t.bond_atoms = D->bond_atoms.data(); t.bond_k = D->bond_k.data(); t.bond_r0 = D->bond_r0.data();
    t.n_angles = na; t.angle_atoms = D->angle_atoms.data(); t.angle_k = D->angle_k.data();
    t.angle_theta0 = D->angle_theta0.data(); t.ub_k = D->ub_k.data(); t.ub_s0 = D->ub_s0.data();
    t.n_dihedrals = (int)D->dihedral_k.size(); t.dihedral_atoms = D->dihedral_atoms.data();
    t.dihedral_k = D->dihedral_k.data(); t.dihedral_n = D->dihedral_n.data(); t.dihedral_delta = D->dihedral_delta.data();
    t.n_impropers = ni; t.improper_atoms = D->improper_atoms.data(); t.improper_k = D->improper_k.data();
    t.improper_psi0 = D->improper_psi0.data();
    t.n_constraints = (int)D->constraint_length.size();
    t.constraint_atoms = D->constraint_atoms.data(); t.constraint_length = D->constraint_length.data();
    fz_params& p = D->par;
    for (int d = 0; d < 3; d++) { p.box[d] = box[d]; p.pme_grid[d] = K[d]; }
    p.cutoff = cutoff; p.cuton = cuton; p.pairlist = pairlist; p.beta = kappa;
    p.pme_order = 4; p.dt_ps = dt; p.shake_tol = shaketol;
    fz_run_settings& s = D->set;
    s.temperature = T; s.dynasteps = dynasteps; s.minimizesteps = minimize; s.startstep = startstep;
    s.tempreset_every = tempreset; s.output_every = outputTiming; s.seed = (uint64_t)seed;
    s.shake = shake; s.print_to_stdout = printstdout;
    snprintf(s.output_prefix, sizeof s.output_prefix, "%s", outname.c_str());
    s.cgfac = cgfac; s.sdfac = sdfac; s.cgrate = cgrate; s.sdrate = sdrate;
    s.checkpoint_every = ckpt_every; s.traj_every = traj_every > 0 ? traj_every : outputTiming;
    s.print_level = printlevel; s.checkpoint_binary = ckpt_binary;
    s.pmass = pmass; s.pref = pref;
    {  // <outputtrajectory> without the extension after its last '/' (src/fenzi.cu:150-158)
      std::string pfx = outname.empty() ? std::string("fenzi") : outname;
      const size_t dot = pfx.rfind('.'), sl = pfx.rfind('/');
      if (dot != std::string::npos && (sl == std::string::npos || dot > sl)) pfx.resize(dot);
      snprintf(s.file_prefix, sizeof s.file_prefix, "%s", pfx.c_str());
    }
    *out = guard.release();
    return FZ_OK;
  } catch (...) {
    return set_error(FZ_ERR_PARSE, "exception while reading %s", in_path ? in_path : "(null)");
  }
}

void fz_deck_destroy(fz_deck* d) { delete d; }
const fz_topology* fz_deck_topology(const fz_deck* d) { return d ? &d->topo : nullptr; }
const fz_params* fz_deck_params(const fz_deck* d) { return d ? &d->par : nullptr; }
const double* fz_deck_coordinates(const fz_deck* d) { return d ? d->xyz.data() : nullptr; }
int fz_deck_settings(const fz_deck* d, fz_run_settings* out) {
  if (!d || !out) return set_error(FZ_ERR_ARG, "null argument");
  *out = d->set;
  return FZ_OK;
}
int fz_deck_psf_counts(const fz_deck* d, int64_t* counts) {
  if (!d || !counts) return set_error(FZ_ERR_ARG, "null argument");
  for (int k = 0; k < 4; k++) counts[k] = d->psf_counts[k];
  return FZ_OK;
}

// consharm groups by CRD segment name, compared case-insensitively
// (cicompare, src/prmread.cu:366-378)
int fz_deck_com_restraint(const fz_deck* d, int32_t* group, double* dist, double* fc) {
  if (!d || !group || !dist || !fc) return set_error(FZ_ERR_ARG, "null argument");
  if (!d->consharm) return 0;
  const std::string s0 = lower(d->consharm_seg[0]), s1 = lower(d->consharm_seg[1]);
  const size_t n = d->charge.size();
  for (size_t i = 0; i < n; i++) {
    const std::string seg = i < d->crd_seg.size() ? lower(d->crd_seg[i]) : std::string();
    group[i] = -1;
    if (!seg.empty() && seg == s0) group[i] = 0;
    if (!seg.empty() && seg == s1) group[i] = 1;
  }
  for (int k = 0; k < 3; k++) dist[k] = d->consharm_dist[k];
  *fc = d->consharm_fc;
  return 1;
}

// consfix segments by PSF segment name (src/prmread.cu:1675-1697)
int fz_deck_fixed_atoms(const fz_deck* d, uint8_t* fixed) {
  if (!d || !fixed) return set_error(FZ_ERR_ARG, "null argument");
  const size_t n = d->charge.size();
  int nf = 0;
  for (size_t i = 0; i < n; i++) {
    fixed[i] = 0;
    const std::string seg = lower(d->psf_seg[i]);
    for (const auto& f : d->consfix_seg)
      if (seg == lower(f)) fixed[i] = 1;
    nf += fixed[i];
  }
  return nf;
}

// Energy table (src/evalprops.cu:328-343 header, :248-267 row; UREY_BRADLEY,
// IMPROPER and PME_CALC columns on).  "Potential" is the sum of the terms; the
// reference's per-atom .w sum double-counts 1-4 pairs.
int fz_format_energy_header(char* buf, size_t len) {
  return snprintf(buf, len, "%-9s %-10s %-10s %-11s %-11s %-11s %-8s %-11s %-11s %-13s %-11s %-13s %-13s %-13s %-13s %-13s\n",
                  "Step", "Time (ps)", "Temp (K)", "bond", "angle", "urey-b", "impr", "dihed", "VDW", "elec", "PME",
                  "Ewexcl", "Ewself", "sum PE", "Kinetic", "Potential");
}

int fz_format_energy_row(const fz_energies* e, double dt_ps, char* buf, size_t len) {
  if (!e || !buf) return set_error(FZ_ERR_ARG, "null argument");
  return snprintf(buf, len,
                  "%-9lld %-10.4f %-10.4f %-11.4f %-11.4f %-11.4f %-8.4f %-11.4f %-11.4f %-13.4f %-11.4f %-13.4f %-13.4f "
                  "%-13.4f %-13.4f %-13.4f\n",
                  (long long)e->step, e->step * dt_ps, e->temperature, e->bond, e->angle, e->urey_bradley, e->improper,
                  e->dihedral, e->vdw, e->elec, e->pme, e->ewald_excl, e->ewald_self, e->potential, e->kinetic,
                  e->potential);
}

}  // extern "C"
