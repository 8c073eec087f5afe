// fenzi_main.cpp -- the `fenzi -d <device> -f <file.in>` driver over libfenzi.
//
// Replaces main() of the reference (src/fenzi.cu:112-630): read the deck, set
// up the system, assign Maxwell velocities, run the dynamics loop with NVT
// velocity reassignment every tempresetTiming steps and the energy table
// every outputTiming steps, then report the wall-clock rate (the reference
// times with clock(), i.e. CPU time, src/fenzi.cu:441,557).
// Extra flag: -p fp64|mixed selects the precision form (default fp64).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fenzi.h"

static int die(const char* what) {
  fprintf(stderr, "fenzi: %s: %s\n", what, fz_last_error());
  return 1;
}

int main(int argc, char** argv) {
  int device = 0;
  const char* in = nullptr;
  int prec = FZ_FP64;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-d") && i + 1 < argc) device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "-f") && i + 1 < argc) in = argv[++i];
    else if (!strcmp(argv[i], "-p") && i + 1 < argc) {
      std::string p = argv[++i];
      prec = p == "mixed" ? FZ_MIXED : FZ_FP64;
    } else {
      fprintf(stderr, "usage: fenzi -d <device> -f <file.in> [-p fp64|mixed]\n");
      return 1;
    }
  }
  if (!in) {
    fprintf(stderr, "usage: fenzi -d <device> -f <file.in> [-p fp64|mixed]\n");
    return 1;
  }
  fz_deck* deck = nullptr;
  if (fz_deck_load(in, &deck)) return die("reading input");
  fz_run_settings rs;
  fz_deck_settings(deck, &rs);
  fz_ctx* ctx = nullptr;
  if (fz_ctx_create(device, &ctx)) return die("device");
  fz_system* sys = nullptr;
  const fz_params* par = fz_deck_params(deck);
  fz_params run_par = *par;
  const int64_t n_atoms = fz_deck_topology(deck)->n_atoms;
  // restart from <prefix>.rst when it exists (src/fenzi.cu:161-181)
  const std::string rst = std::string(rs.file_prefix) + ".rst";
  bool restart = false;
  std::vector<double> rx((size_t)3 * n_atoms), rv((size_t)3 * n_atoms);
  if (FILE* probe = fopen(rst.c_str(), "rb")) {
    fclose(probe);
    int64_t rstep = 0;
    double rseed = 0, rbox[3];
    if (fz_restart_read(rst.c_str(), rs.checkpoint_binary, n_atoms, &rstep, &rseed, rbox, rx.data(), rv.data()))
      return die("reading the restart file");
    rs.startstep = rstep;
    restart = true;
    // a barostat run resumes at the box it had reached
    if (rs.pmass > 0)
      for (int k = 0; k < 3; k++) run_par.box[k] = rbox[k];
  }
  if (fz_system_create(ctx, fz_deck_topology(deck), &run_par, prec, &sys)) return die("system setup");
  if (fz_set_state(sys, restart ? rx.data() : fz_deck_coordinates(deck), restart ? rv.data() : nullptr))
    return die("initial state");
  {  // consfix segments and the consharm restraint (src/prmread.cu:674-722)
    std::vector<uint8_t> fixed((size_t)n_atoms);
    const int nf = fz_deck_fixed_atoms(deck, fixed.data());
    if (nf < 0) return die("fixed atoms");
    if (nf > 0 && fz_set_fixed_atoms(sys, fixed.data())) return die("fixed atoms");
    std::vector<int32_t> group((size_t)n_atoms);
    double dist[3], fc = 0.0;
    const int on = fz_deck_com_restraint(deck, group.data(), dist, &fc);
    if (on < 0) return die("restraint");
    if (on > 0 && fz_set_com_restraint(sys, group.data(), dist, fc)) return die("restraint");
  }
  if (rs.minimizesteps > 0) {  // the reference minimises before dynamics (src/fenzi.cu:397-420)
    fz_min_params mp;
    mp.method = 1;
    mp.max_step = 0.05;
    mp.cgfac = rs.cgfac; mp.sdfac = rs.sdfac; mp.cgrate = rs.cgrate; mp.sdrate = rs.sdrate;
    if (fz_minimize_ex(sys, (int)rs.minimizesteps, &mp)) return die("minimization");
  }
  // Andersen barostat from the deck's pmass / pref (the box of the restart file,
  // when resuming, is the one the state was read at)
  if (rs.pmass > 0 && fz_set_barostat(sys, rs.pref, rs.pmass)) return die("barostat");
  // velocity draws are counted from the run's first step, so a restarted run
  // continues the same sequence
  uint64_t counter = rs.tempreset_every > 0 ? (uint64_t)(rs.startstep / rs.tempreset_every) + 1 : 1;
  if (!restart && fz_reassign_velocities(sys, rs.temperature, rs.seed, 0)) return die("velocities");
  FILE* out = stdout;
  std::string outpath;
  if (!rs.print_to_stdout) {
    outpath = std::string(rs.file_prefix) + "_dev" + std::to_string(device) + ".out";
    out = fopen(outpath.c_str(), restart ? "a" : "w");
    if (!out) { fprintf(stderr, "fenzi: cannot open %s\n", outpath.c_str()); return 1; }
  }
  const double dt = par->dt_ps;
  fz_dcd* dcd = nullptr;
  if (rs.print_level > 1) {  // trajectory (src/fenzi.cu:280-369)
    const std::string path = std::string(rs.file_prefix) + "_dev" + std::to_string(device) + ".dcd";
    if (fz_dcd_open(path.c_str(), n_atoms, rs.dynasteps / rs.traj_every, rs.traj_every, rs.traj_every, rs.dynasteps,
                    dt, &dcd))
      return die("trajectory");
  }
  char buf[1024];
  fz_format_energy_header(buf, sizeof buf);
  fputs(buf, out);
  fz_energies e;
  if (fz_compute_forces(sys, nullptr, &e)) return die("forces");
  e.step = rs.startstep;
  fz_format_energy_row(&e, dt, buf, sizeof buf);
  fputs(buf, out);
  const long every = rs.output_every > 0 ? rs.output_every : 100;
  long step = rs.startstep;
  // dynasteps is the last step number, also after a restart (src/fenzi.cu:442)
  const long end = rs.dynasteps > rs.startstep ? rs.dynasteps : rs.startstep;
  auto t0 = std::chrono::steady_clock::now();
  auto until = [&](long period) { return period > 0 ? period - step % period : end - step; };
  while (step < end) {
    long chunk = std::min({until(every), until(rs.tempreset_every), until(rs.checkpoint_every),
                           dcd ? until(rs.traj_every) : end - step, end - step});
    if (fz_step(sys, (int)chunk)) return die("dynamics");
    step += chunk;
    if (step % every == 0 || step == end) {
      fz_get_energies(sys, &e);
      e.step = step;
      fz_format_energy_row(&e, dt, buf, sizeof buf);
      fputs(buf, out);
      fflush(out);
    }
    const bool ckpt = rs.print_level > 0 && rs.checkpoint_every > 0 && step % rs.checkpoint_every == 0;
    const bool frame = dcd && step % rs.traj_every == 0;
    if (ckpt || frame) {
      if (fz_get_state(sys, rx.data(), rv.data())) return die("state");
      double box_now[3];  // the barostat's current box (the deck's box otherwise)
      if (fz_get_barostat(sys, box_now, nullptr)) return die("box");
      if (ckpt && fz_restart_write(rst.c_str(), rs.checkpoint_binary, n_atoms, step, (double)rs.seed, box_now,
                                   rx.data(), rv.data()))
        return die("checkpoint");
      if (frame && fz_dcd_write_frame(dcd, rx.data(), box_now)) return die("trajectory");
    }
    if (rs.tempreset_every > 0 && step % rs.tempreset_every == 0 && step < end)
      if (fz_reassign_velocities(sys, rs.temperature, rs.seed, counter++)) return die("velocities");
  }
  if (dcd) fz_dcd_close(dcd);
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  long nsteps = end - rs.startstep;
  if (nsteps > 0) {
    double per = secs / nsteps;
    fprintf(out, "Avg time/step (wall clock): %.6f ms\n", per * 1e3);
    fprintf(out, "Ns per day: %.4f\n", 86400.0 / per * dt * 1e-3);
  }
  if (out != stdout) fclose(out);
  fz_system_destroy(sys);
  fz_ctx_destroy(ctx);
  fz_deck_destroy(deck);
  return 0;
}
