// fz_io.cpp -- checkpoint (.rst) and CHARMM trajectory (.dcd) files.
//
// Restart files keep the reference's two layouts (SaveCheckpointBinary /
// SaveCheckpointAscii and their loaders, src/prmwrite.cu:31-252), so that a
// run can be resumed from a file the reference wrote and the reverse: atom
// count, step, seed, box, three zero box velocities, then per atom position
// and velocity, in single precision (binary: native int32 / float64 seed /
// float32; text: the reference's printf formats).  Trajectories are CHARMM
// DCD files with Fortran record framing (charmm_trj_write_* in
// src/io/charmmio.c:50-175, ffwrite in src/io/ffile.c:155-203): a "CORD"
// header, one title, the atom count, then per frame the unit-cell record and
// the x, y, z records in single precision.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fz_engine.h"

using namespace fz;

struct fz_dcd {
  FILE* f = nullptr;
  int64_t n = 0;
  std::vector<float> buf;
};

namespace {

// one Fortran unformatted record: length, payload pieces, length
bool record(FILE* f, std::initializer_list<std::pair<const void*, size_t>> parts) {
  uint32_t len = 0;
  for (auto& p : parts) len += (uint32_t)p.second;
  if (fwrite(&len, 4, 1, f) != 1) return false;
  for (auto& p : parts)
    if (p.second && fwrite(p.first, 1, p.second, f) != p.second) return false;
  return fwrite(&len, 4, 1, f) == 1;
}

}  // namespace

extern "C" {

int fz_restart_write(const char* path, int32_t binary, int64_t n_atoms, int64_t step, double seed, const double* box,
                     const double* xyz, const double* vel) {
  if (!path || !box || !xyz || !vel || n_atoms < 0 || n_atoms > INT32_MAX || step > INT32_MAX)
    return set_error(FZ_ERR_ARG, "invalid argument");
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return set_error(FZ_ERR_ARG, "cannot open %s for writing", path);
  bool ok = true;
  if (binary) {
    const int32_t n32 = (int32_t)n_atoms, s32 = (int32_t)step;
    const float b[3] = {(float)box[0], (float)box[1], (float)box[2]}, z[3] = {0.0f, 0.0f, 0.0f};
    ok = fwrite(&n32, 4, 1, f) == 1 && fwrite(&s32, 4, 1, f) == 1 && fwrite(&seed, 8, 1, f) == 1 &&
         fwrite(b, 4, 3, f) == 3 && fwrite(z, 4, 3, f) == 3;
    std::vector<float> rec((size_t)6 * n_atoms);
    for (int64_t i = 0; i < n_atoms; i++)
      for (int k = 0; k < 3; k++) {
        rec[6 * i + k] = (float)xyz[3 * i + k];
        rec[6 * i + 3 + k] = (float)vel[3 * i + k];
      }
    ok = ok && fwrite(rec.data(), 4, rec.size(), f) == rec.size();
  } else {
    ok = fprintf(f, "%12d \n%12d \n%le \n", (int)n_atoms, (int)step, seed) > 0 &&
         fprintf(f, "%12f %12f %12f %12f %12f %12f\n", (float)box[0], (float)box[1], (float)box[2], 0.0f, 0.0f, 0.0f) > 0;
    for (int64_t i = 0; ok && i < n_atoms; i++)
      ok = fprintf(f, "%10f %10f %10f %10f %10f %10f\n", (float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2],
                   (float)vel[3 * i], (float)vel[3 * i + 1], (float)vel[3 * i + 2]) > 0;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? FZ_OK : set_error(FZ_ERR_ARG, "writing %s failed", path);
}

int fz_restart_read(const char* path, int32_t binary, int64_t n_atoms, int64_t* step, double* seed, double* box,
                    double* xyz, double* vel) {
  if (!path || n_atoms < 0) return set_error(FZ_ERR_ARG, "invalid argument");
  FILE* f = fopen(path, binary ? "rb" : "r");
  if (!f) return set_error(FZ_ERR_ARG, "restart file %s not found", path);
  int32_t n32 = 0, s32 = 0;
  double sd = 0.0;
  float b[3] = {0, 0, 0}, z[3] = {0, 0, 0};
  bool ok;
  std::vector<float> rec;
  if (binary) {
    ok = fread(&n32, 4, 1, f) == 1 && fread(&s32, 4, 1, f) == 1 && fread(&sd, 8, 1, f) == 1 && fread(b, 4, 3, f) == 3 &&
         fread(z, 4, 3, f) == 3;
    if (ok && n32 == n_atoms) {
      rec.resize((size_t)6 * n_atoms);
      ok = fread(rec.data(), 4, rec.size(), f) == rec.size();
    }
  } else {
    ok = fscanf(f, "%d %d %lf %f %f %f %f %f %f", &n32, &s32, &sd, &b[0], &b[1], &b[2], &z[0], &z[1], &z[2]) == 9;
    if (ok && n32 == n_atoms) {
      rec.resize((size_t)6 * n_atoms);
      for (size_t k = 0; ok && k < rec.size(); k++) ok = fscanf(f, "%f", &rec[k]) == 1;
    }
  }
  fclose(f);
  if (!ok) return set_error(FZ_ERR_PARSE, "unexpected end of restart file %s", path);
  if (n32 != n_atoms) return set_error(FZ_ERR_PARSE, "restart file %s holds %d atoms, the system %lld", path, n32, (long long)n_atoms);
  if (step) *step = s32;
  if (seed) *seed = sd;
  if (box)
    for (int k = 0; k < 3; k++) box[k] = b[k];
  for (int64_t i = 0; i < n_atoms; i++)
    for (int k = 0; k < 3; k++) {
      if (xyz) xyz[3 * i + k] = rec[6 * i + k];
      if (vel) vel[3 * i + k] = rec[6 * i + 3 + k];
    }
  return FZ_OK;
}

int fz_dcd_open(const char* path, int64_t n_atoms, int64_t nframes, int64_t first_step, int64_t every, int64_t nsteps,
                double dt_ps, fz_dcd** out) {
  if (!path || !out || n_atoms <= 0 || n_atoms > INT32_MAX) return set_error(FZ_ERR_ARG, "invalid argument");
  FILE* f = fopen(path, "wb");
  if (!f) return set_error(FZ_ERR_ARG, "cannot open %s for writing", path);
  int32_t h[20];
  memset(h, 0, sizeof h);
  h[0] = (int32_t)nframes;
  h[1] = (int32_t)first_step;
  h[2] = (int32_t)(every > 0 ? every : 1);
  h[3] = (int32_t)nsteps;
  h[7] = 1;  // degrees-of-freedom field as the reference fills it
  h[8] = 0;  // no fixed atoms
  const float dt_akma = (float)(dt_ps / 4.888821e-2);
  memcpy(&h[9], &dt_akma, 4);
  h[10] = 1;  // unit-cell record present
  h[19] = 34; // CHARMM version
  char title[80];
  memset(title, ' ', sizeof title);
  const char* t = "Generated by fenzi (MI355X build)";
  memcpy(title, t, strlen(t));
  const int32_t ntitle = 1, na = (int32_t)n_atoms;
  bool ok = record(f, {{"CORD", 4}, {h, sizeof h}}) && record(f, {{&ntitle, 4}, {title, 80}}) && record(f, {{&na, 4}});
  if (!ok) {
    fclose(f);
    return set_error(FZ_ERR_ARG, "writing %s failed", path);
  }
  auto* d = new fz_dcd();
  d->f = f;
  d->n = n_atoms;
  d->buf.resize((size_t)n_atoms);
  *out = d;
  return FZ_OK;
}

int fz_dcd_write_frame(fz_dcd* d, const double* xyz, const double* box) {
  if (!d || !xyz || !box) return set_error(FZ_ERR_ARG, "invalid argument");
  const double cell[6] = {box[0], 0.0, box[1], 0.0, 0.0, box[2]};
  bool ok = record(d->f, {{cell, sizeof cell}});
  for (int k = 0; ok && k < 3; k++) {
    for (int64_t i = 0; i < d->n; i++) d->buf[i] = (float)xyz[3 * i + k];
    ok = record(d->f, {{d->buf.data(), sizeof(float) * d->buf.size()}});
  }
  fflush(d->f);
  return ok ? FZ_OK : set_error(FZ_ERR_ARG, "writing a trajectory frame failed");
}

int fz_dcd_close(fz_dcd* d) {
  if (!d) return FZ_OK;
  const int r = fclose(d->f);
  delete d;
  return r == 0 ? FZ_OK : set_error(FZ_ERR_ARG, "closing the trajectory failed");
}

}  // extern "C"
