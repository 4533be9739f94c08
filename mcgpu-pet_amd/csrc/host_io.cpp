// host_io.cpp — the file-format boundary of the photon-transport path.
//
// Reads the reference's MCGPU-PET.in, .vox[.gz] and .mcgpu[.gz] material files and
// writes its output files. The parsing rules (section markers found with strstr,
// positional values, '#' comments, the table arithmetic done in double and stored in
// float) follow MCGPU-PET.cu; each function cites the lines it replaces. Errors are
// returned as mcpet_status codes instead of exit().
#include "mcpet.h"
#include "host_internal.h"

#include <zlib.h>
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

mcpet_status fail(mcpet_status st, char* err, size_t errlen, const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  if (err && errlen) {
    strncpy(err, buf, errlen - 1);
    err[errlen - 1] = '\0';
  }
  return st;
}

// Line source over a FILE* or a gzFile (the reference reads .in with fgets and
// .vox/.mcgpu with gzgets; both cap a line at 250 bytes including the newline).
struct LineReader {
  FILE* f = nullptr;
  gzFile g = nullptr;
  bool eof_flag = false;
  explicit LineReader(FILE* fp) : f(fp) {}
  explicit LineReader(gzFile gp) : g(gp) {}
  // Returns false at end of file. buf always NUL-terminated.
  bool gets(char* buf, int n)
  {
    char* r = f ? fgets(buf, n, f) : gzgets(g, buf, n);
    if (!r) {
      eof_flag = true;
      buf[0] = '\0';
      return false;
    }
    return true;
  }
  bool eof() const { return f ? feof(f) != 0 : gzeof(g) != 0; }
};

// fgets_trimmed (MCGPU-PET.cu:1320-1350): next line that is not empty or comment-only,
// with leading blanks and the trailing comment removed. Returns false at EOF
// (out is then empty).
bool gets_trimmed(LineReader& rd, char* out, int n)
{
  char line[MCPET_NAME_LEN + 2];
  out[0] = '\0';
  for (;;) {
    if (!rd.gets(line, n)) return false;
    int a = 0, b = 0;
    while (line[a] == ' ') a++;
    while (line[a] != '\n' && line[a] != '#' && line[a] != '\0' && line[a] != '\r') out[b++] = line[a++];
    out[b] = '\0';
    if (out[0] != '\0') return true;
  }
}

// trim_name (MCGPU-PET.cu:1292-1311): first blank-free token, stops at ' ' or '#'.
void trim_name(const char* in, char* name)
{
  int a = 0, b = 0;
  while (in[a] == ' ') a++;
  while (in[a] != ' ' && in[a] != '#' && in[a] != '\0' && in[a] != '\t' && in[a] != '\r' && in[a] != '\n')
    name[b++] = in[a++];
  name[b] = '\0';
}

// Skip raw lines until one contains marker (the do/while strstr loops of read_input).
bool find_section(LineReader& rd, const char* marker, char* line, const char* bad_marker = nullptr,
                  bool* saw_bad = nullptr)
{
  for (;;) {
    if (!rd.gets(line, MCPET_NAME_LEN)) return false;
    if (bad_marker && strstr(line, bad_marker)) {
      if (saw_bad) *saw_bad = true;
      return false;
    }
    if (strstr(line, marker)) return true;
  }
}

bool starts_yes(const char* s) { return !strncmp(s, "YE", 2) || !strncmp(s, "Ye", 2) || !strncmp(s, "ye", 2); }
bool starts_no(const char* s) { return !strncmp(s, "NO", 2) || !strncmp(s, "No", 2) || !strncmp(s, "no", 2); }

}  // namespace

const char* mcpet_host_last_error() { return g_last_error.c_str(); }
void mcpet_host_set_error(const char* msg) { g_last_error = msg; }

extern "C" void mcpet_sinogram_sizes(mcpet_config* c)
{
  // MCGPU-PET.cu:1176-1189 (integer division, as in the reference).
  c->nseg = 2 * (c->mrd / c->span) + 1;
  int nsinos = c->nseg * c->nzs;
  for (int aaa = 1; aaa <= c->nseg; aaa++) {
    if (aaa > 1) nsinos -= c->span + 1;
    if (aaa > 3) nsinos -= 2 * ((aaa - 2) / 2) * c->span;
  }
  c->nsinos = nsinos;
  c->nbins = c->nrad * c->nangles * c->nsinos;
  c->nvoxs = c->res * c->res * c->nzs;
}

// Ring sum of every sinogram: segment iseg starts at ofseg(iseg) and its first sinogram holds
// ring sum ofz(incl), incl = (iseg + 1) / 2 (the offsets of the binning, K:1749-1788).
extern "C" mcpet_status mcpet_sinogram_planes(const mcpet_config* cfg, int32_t* plane)
{
  if (!cfg || !plane || cfg->span < 1 || cfg->nseg < 1 || cfg->nsinos < 1) {
    mcpet_host_set_error("mcpet_sinogram_planes: bad arguments (config sizes not derived?)");
    return MCPET_E_ARG;
  }
  const int span = cfg->span;
  auto ofseg = [&](int iseg) {
    int o = iseg * cfg->nzs;
    for (int kka = 2; kka <= iseg; kka++) {
      o -= span + 1;
      if (kka > 3) o -= ((kka - 2) / 2) * 2 * span;
    }
    return o;
  };
  for (int iseg = 0; iseg < cfg->nseg; iseg++) {
    const int incl = (iseg + 1) / 2;
    const int ofz = incl > 0 ? (span + 1) / 2 + (incl - 1) * span : 0;
    const int b = ofseg(iseg), e = std::min(ofseg(iseg + 1), cfg->nsinos);
    for (int izm = b; izm < e; izm++) plane[izm] = izm - b + ofz;
  }
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_read_input(const char* path, mcpet_config* c, char* err, size_t errlen)
{
  if (!path || !c) return fail(MCPET_E_ARG, err, errlen, "mcpet_read_input: null argument");
  memset(c, 0, sizeof *c);
  FILE* fp = fopen(path, "r");
  if (!fp) return fail(MCPET_E_IO, err, errlen, "read_input: input file not found or not readable: '%s'", path);
  LineReader rd(fp);
  char line[MCPET_NAME_LEN + 2], t[MCPET_NAME_LEN + 2];
  struct Closer { FILE* f; ~Closer() { fclose(f); } } closer{fp};

#define NEED_SECTION(marker)                                                                         \
  if (!find_section(rd, marker, line))                                                               \
    return fail(MCPET_E_FORMAT, err, errlen, "read_input: file does not contain the string '%s'", marker);
#define NEXT_VALUE(what)                                                                             \
  if (!gets_trimmed(rd, t, MCPET_NAME_LEN))                                                          \
    return fail(MCPET_E_FORMAT, err, errlen, "read_input: unexpected end of file reading %s", what);
#define SCAN1(fmt, ptr, what)                                                                        \
  NEXT_VALUE(what)                                                                                   \
  if (sscanf(t, fmt, ptr) != 1) return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad %s: '%s'", what, t);

  // [SECTION SIMULATION CONFIG v.2016-07-05]  (C:906-956)
  NEED_SECTION("SECTION SIMULATION CONFIG v.2016-07-05");
  SCAN1("%d", &c->seed, "random seed");
  if (c->seed == 0) {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    c->seed = (int)((tv.tv_sec / 1000) + (tv.tv_usec * 1000));
  }
  SCAN1("%d", &c->gpu_id, "GPU number");
  SCAN1("%d", &c->threads_per_block, "threads per block");
  SCAN1("%f", &c->fact_object, "density factor");
  if (c->threads_per_block < 1) return fail(MCPET_E_RANGE, err, errlen, "read_input: threads per block must be >= 1");

  // [SECTION SOURCE PET SCAN v.2017-03-14]  (C:963-995)
  NEED_SECTION("SECTION SOURCE PET SCAN v.2017-03-14");
  double T = 0.0;
  SCAN1("%lf", &T, "acquisition time");
  if (!(T >= 0.0) || T * 1.0e12 > 1.8e19) return fail(MCPET_E_RANGE, err, errlen, "read_input: bad acquisition time %g", T);
  c->acquisition_time_ps = (uint64_t)(T * 1.0e12 + 0.5);
  SCAN1("%f", &c->mean_life, "isotope mean life");
  for (int ii = 0; ii < MCPET_MAX_MATERIALS; ii++) {
    int mat = 0;
    float act = 0.0f;
    NEXT_VALUE("activity table");
    if (sscanf(t, "%d %f", &mat, &act) < 1) return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad activity line '%s'", t);
    if (mat < 1) break;
    if (mat > MCPET_MAX_MATERIALS) return fail(MCPET_E_RANGE, err, errlen, "read_input: activity material %d > 15", mat);
    c->activity[mat - 1] = act;
  }

  // [SECTION PHASE SPACE FILE v.2016-07-05]  (C:1000-1026)
  NEED_SECTION("SECTION PHASE SPACE FILE v.2016-07-05");
  NEXT_VALUE("PSF file name");
  trim_name(t, c->psf_file);
  NEXT_VALUE("detector geometry");
  if (sscanf(t, "%f %f %f %f %f", &c->detector.center[0], &c->detector.center[1], &c->detector.center[2],
             &c->detector.height, &c->detector.radius) != 5)
    return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad detector line '%s'", t);
  SCAN1("%d", &c->detector.psf_size, "PSF size");
  SCAN1("%d", &c->detector.tally_type, "tally type");
  if (c->detector.tally_type < 0 || c->detector.tally_type > 2) c->detector.tally_type = 0;
  SCAN1("%d", &c->detector.tally_psf_sinogram, "PSF/sinogram flag");
  if (c->detector.tally_psf_sinogram < 0 || c->detector.tally_psf_sinogram > 2) c->detector.tally_psf_sinogram = 0;

  // [SECTION DOSE DEPOSITION v.2012-12-12]  (C:1032-1117)
  {
    bool old = false;
    if (!find_section(rd, "SECTION DOSE DEPOSITION v.2012-12-12", line, "SECTION DOSE DEPOSITION v.2011-02-18", &old)) {
      if (old) return fail(MCPET_E_FORMAT, err, errlen, "read_input: please update the input file to MC-GPU v1.3 dose section");
      return fail(MCPET_E_FORMAT, err, errlen, "read_input: file does not contain 'SECTION DOSE DEPOSITION v.2012-12-12'");
    }
  }
  NEXT_VALUE("material dose flag");
  if (starts_yes(t)) c->material_dose = 1;
  else if (starts_no(t)) c->material_dose = 0;
  else return fail(MCPET_E_FORMAT, err, errlen, "read_input: answer YES or NO for the material dose tally: '%s'", t);
  NEXT_VALUE("voxel dose flag");
  if (starts_yes(t)) {
    c->voxel_dose = 1;
    NEXT_VALUE("dose file name");
    trim_name(t, c->dose_file);
    short v0, v1;
    NEXT_VALUE("dose ROI x");
    if (sscanf(t, "%hd %hd", &v0, &v1) != 2) return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad ROI line '%s'", t);
    c->roi_x_min = v0 - 1; c->roi_x_max = v1 - 1;
    NEXT_VALUE("dose ROI y");
    if (sscanf(t, "%hd %hd", &v0, &v1) != 2) return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad ROI line '%s'", t);
    c->roi_y_min = v0 - 1; c->roi_y_max = v1 - 1;
    NEXT_VALUE("dose ROI z");
    if (sscanf(t, "%hd %hd", &v0, &v1) != 2) return fail(MCPET_E_FORMAT, err, errlen, "read_input: bad ROI line '%s'", t);
    c->roi_z_min = v0 - 1; c->roi_z_max = v1 - 1;
    if (c->roi_x_min > c->roi_x_max || c->roi_y_min > c->roi_y_max || c->roi_z_min > c->roi_z_max ||
        c->roi_x_min < 0 || c->roi_y_min < 0 || c->roi_z_min < 0)
      return fail(MCPET_E_RANGE, err, errlen, "read_input: the dose region of interest is not valid");
  } else if (starts_no(t)) {
    c->voxel_dose = 0;
    c->roi_x_min = c->roi_y_min = c->roi_z_min = 32500;
    c->roi_x_max = c->roi_y_max = c->roi_z_max = -32500;
  } else {
    return fail(MCPET_E_FORMAT, err, errlen, "read_input: answer YES or NO for the voxel dose tally: '%s'", t);
  }

  // [SECTION ENERGY PARAMETERS v.2019-04-25]  (C:1120-1139)
  NEED_SECTION("SECTION ENERGY PARAMETERS v.2019-04-25");
  SCAN1("%f", &c->e_resol, "energy resolution");
  SCAN1("%f", &c->e_low, "energy window low");
  SCAN1("%f", &c->e_high, "energy window high");

  // [SECTION SINOGRAM PARAMETERS v.2019-04-25]  (C:1141-1189)
  NEED_SECTION("SECTION SINOGRAM PARAMETERS v.2019-04-25");
  SCAN1("%f", &c->fovz, "axial FOV");
  SCAN1("%d", &c->nrows, "number of rows");
  SCAN1("%d", &c->ncrystals, "number of crystals");
  SCAN1("%d", &c->nangles, "number of angular bins");
  SCAN1("%d", &c->nrad, "number of radial bins");
  SCAN1("%d", &c->nzs, "number of z slices");
  SCAN1("%d", &c->res, "image resolution");
  SCAN1("%d", &c->ne, "number of energy bins");
  SCAN1("%d", &c->mrd, "maximum ring difference");
  SCAN1("%d", &c->span, "span");
  if (c->span < 1 || c->nrows < 1 || c->ncrystals < 2 || c->nangles < 1 || c->nrad < 1 || c->nzs < 1 || c->ne < 0 ||
      c->mrd < 0)
    return fail(MCPET_E_RANGE, err, errlen, "read_input: invalid sinogram parameters");
  mcpet_sinogram_sizes(c);
  if ((int64_t)c->nrad * c->nangles * (int64_t)c->nsinos > 0x7fffffffLL || c->nsinos < 1)
    return fail(MCPET_E_RANGE, err, errlen, "read_input: sinogram size out of range");

  // [SECTION VOXELIZED GEOMETRY FILE v.2009-11-30]  (C:1192-1204)
  NEED_SECTION("SECTION VOXELIZED GEOMETRY FILE v.2009-11-30");
  NEXT_VALUE("voxel file name");
  trim_name(t, c->voxel_file);

  // [SECTION MATERIAL FILE LIST]  (C:1206-1225)
  NEED_SECTION("SECTION MATERIAL");
  for (int i = 0; i < MCPET_MAX_MATERIALS; i++) {
    if (!gets_trimmed(rd, t, MCPET_NAME_LEN)) c->material_files[i][0] = '\0';
    else trim_name(t, c->material_files[i]);
  }
  c->bg_mode = 0;
#undef NEED_SECTION
#undef NEXT_VALUE
#undef SCAN1
  return MCPET_OK;
}

// ------------------------------------------------------------------------------- voxels

static mcpet_status finish_phantom_header(mcpet_phantom* ph, mcpet_config* cfg)
{
  ph->size_bbox[0] = ph->nx * ph->voxel_size[0];
  ph->size_bbox[1] = ph->ny * ph->voxel_size[1];
  ph->size_bbox[2] = ph->nz * ph->voxel_size[2];
  ph->inv_voxel_size[0] = 1.0f / ph->voxel_size[0];
  ph->inv_voxel_size[1] = 1.0f / ph->voxel_size[1];
  ph->inv_voxel_size[2] = 1.0f / ph->voxel_size[2];
  for (int k = 0; k < MCPET_MAX_MATERIALS; k++) ph->density_max[k] = -999.0f;
  if (cfg) {
    // C:1444-1452: shrink the ROI maxima to the geometry.
    if (cfg->roi_x_max + 1 > ph->nx) cfg->roi_x_max = (int16_t)(ph->nx - 1);
    if (cfg->roi_y_max + 1 > ph->ny) cfg->roi_y_max = (int16_t)(ph->ny - 1);
    if (cfg->roi_z_max + 1 > ph->nz) cfg->roi_z_max = (int16_t)(ph->nz - 1);
  }
  return MCPET_OK;
}

static mcpet_status store_voxel(mcpet_phantom* ph, int64_t v, int mat, float dens, float act, char* err, size_t errlen)
{
  if (mat > MCPET_MAX_MATERIALS)
    return fail(MCPET_E_RANGE, err, errlen, "load_voxels: voxel material number too high: %d (voxel %lld)", mat, (long long)v + 1);
  if (mat < 1)
    return fail(MCPET_E_RANGE, err, errlen, "load_voxels: voxel material number can not be zero or negative: %d (voxel %lld)", mat, (long long)v + 1);
  if (!(dens >= 1.0e-9f))
    return fail(MCPET_E_RANGE, err, errlen, "load_voxels: voxel density can not be 0 or negative: %g (voxel %lld)", (double)dens, (long long)v + 1);
  if (dens > ph->density_max[mat - 1]) ph->density_max[mat - 1] = dens;
  float* p = ph->voxel_mat_dens + 3 * v;
  p[0] = (float)mat + 0.0001f;   // C:1522
  p[1] = dens;
  p[2] = act;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_load_voxels(const char* path, mcpet_config* cfg, mcpet_phantom* ph, char* err, size_t errlen)
{
  if (!path || !ph) return fail(MCPET_E_ARG, err, errlen, "mcpet_load_voxels: null argument");
  memset(ph, 0, sizeof *ph);
  gzFile g = gzopen(path, "rb");
  if (!g) return fail(MCPET_E_IO, err, errlen, "load_voxels: file %s does not exist", path);
  struct Closer { gzFile g; ~Closer() { gzclose(g); } } closer{g};
  LineReader rd(g);
  char line[MCPET_NAME_LEN + 2];
  if (!find_section(rd, "[SECTION VOXELS", line))
    return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: file does not contain '[SECTION VOXELS HEADER'");
  if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "%d %d %d", &ph->nx, &ph->ny, &ph->nz) != 3)
    return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: bad voxel count line");
  if (!rd.gets(line, MCPET_NAME_LEN) ||
      sscanf(line, "%f %f %f", &ph->voxel_size[0], &ph->voxel_size[1], &ph->voxel_size[2]) != 3)
    return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: bad voxel size line");
  if (ph->nx < 1 || ph->ny < 1 || ph->nz < 1 || (int64_t)ph->nx * ph->ny * ph->nz > 0x7fffffffLL ||
      !(ph->voxel_size[0] > 0) || !(ph->voxel_size[1] > 0) || !(ph->voxel_size[2] > 0))
    return fail(MCPET_E_RANGE, err, errlen, "load_voxels: invalid phantom dimensions");
  if (!find_section(rd, "[END OF VXH SECTION", line))
    return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: file does not contain '[END OF VXH SECTION]'");
  finish_phantom_header(ph, cfg);
  const int64_t nvox = (int64_t)ph->nx * ph->ny * ph->nz;
  ph->voxel_mat_dens = (float*)malloc(sizeof(float) * 3 * (size_t)nvox);
  if (!ph->voxel_mat_dens) return fail(MCPET_E_NOMEM, err, errlen, "load_voxels: out of memory for %lld voxels", (long long)nvox);
  ph->owned = 1;
  const float fact = cfg ? cfg->fact_object : 1.0f;
  for (int64_t v = 0; v < nvox; v++) {
    // C:1490-1491: skip empty and comment lines (first or second character).
    do {
      if (!rd.gets(line, MCPET_NAME_LEN)) {
        mcpet_free_phantom(ph);
        return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: end of file after %lld of %lld voxels", (long long)v, (long long)nvox);
      }
    } while (line[0] == '\n' || line[1] == '\n' || line[0] == '#' || line[1] == '#' || line[0] == '\r');
    int mat = 0;
    float dens = 0.0f, act = 0.0f;
    int items = sscanf(line, "%d %f %f", &mat, &dens, &act);
    if (items != 3) {
      mcpet_free_phantom(ph);
      return fail(MCPET_E_FORMAT, err, errlen, "load_voxels: expecting 3 items per voxel (material, density, activity); read %d at voxel %lld",
                  items, (long long)v + 1);
    }
    dens *= fact;
    mcpet_status st = store_voxel(ph, v, mat, dens, act, err, errlen);
    if (st != MCPET_OK) { mcpet_free_phantom(ph); return st; }
  }
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_phantom_from_arrays(int32_t nx, int32_t ny, int32_t nz, const float voxel_size[3],
                                                  const int32_t* material, const float* density, const float* activity,
                                                  float fact_object, mcpet_phantom* ph, char* err, size_t errlen)
{
  if (!ph || !voxel_size || !material || !density || !activity)
    return fail(MCPET_E_ARG, err, errlen, "mcpet_phantom_from_arrays: null argument");
  memset(ph, 0, sizeof *ph);
  if (nx < 1 || ny < 1 || nz < 1 || (int64_t)nx * ny * nz > 0x7fffffffLL || !(voxel_size[0] > 0) ||
      !(voxel_size[1] > 0) || !(voxel_size[2] > 0))
    return fail(MCPET_E_RANGE, err, errlen, "mcpet_phantom_from_arrays: invalid dimensions");
  ph->nx = nx; ph->ny = ny; ph->nz = nz;
  for (int i = 0; i < 3; i++) ph->voxel_size[i] = voxel_size[i];
  finish_phantom_header(ph, nullptr);
  const int64_t nvox = (int64_t)nx * ny * nz;
  ph->voxel_mat_dens = (float*)malloc(sizeof(float) * 3 * (size_t)nvox);
  if (!ph->voxel_mat_dens) return fail(MCPET_E_NOMEM, err, errlen, "mcpet_phantom_from_arrays: out of memory");
  ph->owned = 1;
  for (int64_t v = 0; v < nvox; v++) {
    mcpet_status st = store_voxel(ph, v, material[v], density[v] * fact_object, activity[v], err, errlen);
    if (st != MCPET_OK) { mcpet_free_phantom(ph); return st; }
  }
  return MCPET_OK;
}

extern "C" void mcpet_free_phantom(mcpet_phantom* ph)
{
  if (!ph) return;
  if (ph->owned) free(ph->voxel_mat_dens);
  ph->voxel_mat_dens = nullptr;
  ph->owned = 0;
}

// ------------------------------------------------------------------------------- materials

extern "C" void mcpet_free_tables(mcpet_tables* t)
{
  if (!t) return;
  if (t->owned) {
    free(t->woodcock);
    free(t->mfp_a);
    free(t->mfp_b);
    free(t->rayleigh);
    free(t->compton);
  }
  t->woodcock = t->mfp_a = t->mfp_b = nullptr;
  t->rayleigh = nullptr;
  t->compton = nullptr;
  t->owned = 0;
}

// Klein-Nishina total cross section per electron, in units of 2 pi r_e^2, at photon energy
// E (eV): k = E / (m_e c^2).
static double klein_nishina(double e)
{
  const double k = e / 510998.95, l = std::log(1.0 + 2.0 * k);
  return (1.0 + k) / (k * k) * (2.0 * (1.0 + k) / (1.0 + 2.0 * k) - l / k) + l / (2.0 * k) -
         (1.0 + 3.0 * k) / ((1.0 + 2.0 * k) * (1.0 + 2.0 * k));
}

// Raw mean-free-path rows (energy, Rayleigh, Compton, photoelectric, total, pmax) of one
// material past its file's last point: the continuation mcpet_load_materials_extended documents.
static void extrapolate_rows(std::vector<double>& raw, int n_file, int n_total, double e0, double delta_e)
{
  const double* L = &raw[6 * (size_t)(n_file - 1)];
  const double EL = L[0];
  int h = (int)std::floor((0.5 * EL - e0) / delta_e + 0.5);
  h = std::max(0, std::min(h, n_file - 2));
  const double* Hh = &raw[6 * (size_t)h];
  const double s_ray = std::log(L[1] / Hh[1]) / std::log(EL / Hh[0]);
  const double s_pho = std::log(L[3] / Hh[3]) / std::log(EL / Hh[0]);
  const double kn_last = klein_nishina(EL);
  for (int i = n_file; i < n_total; i++) {
    double* r = &raw[6 * (size_t)i];
    const double e = e0 + i * delta_e;
    r[0] = e;
    r[1] = L[1] * std::pow(e / EL, s_ray);
    r[2] = L[2] * kn_last / klein_nishina(e);
    r[3] = L[3] * std::pow(e / EL, s_pho);
    r[4] = 1.0 / (1.0 / r[1] + 1.0 / r[2] + 1.0 / r[3]);
    r[5] = L[5];
  }
}

extern "C" mcpet_status mcpet_load_materials(const char names[MCPET_MAX_MATERIALS][MCPET_NAME_LEN],
                                             const float density_max_in[MCPET_MAX_MATERIALS], mcpet_tables* T,
                                             char* err, size_t errlen)
{
  return mcpet_load_materials_extended(names, density_max_in, 0.0, T, err, errlen);
}

extern "C" mcpet_status mcpet_load_materials_extended(const char names[MCPET_MAX_MATERIALS][MCPET_NAME_LEN],
                                                      const float density_max_in[MCPET_MAX_MATERIALS], double e_max,
                                                      mcpet_tables* T, char* err, size_t errlen)
{
  if (!names || !density_max_in || !T) return fail(MCPET_E_ARG, err, errlen, "mcpet_load_materials: null argument");
  memset(T, 0, sizeof *T);
  float density_max[MCPET_MAX_MATERIALS];
  for (int m = 0; m < MCPET_MAX_MATERIALS; m++) {
    density_max[m] = density_max_in[m];
    T->density_nominal[m] = -1.0f;
  }
  T->rayleigh = (mcpet_rayleigh_table*)calloc(1, sizeof(mcpet_rayleigh_table));
  T->compton = (mcpet_compton_table*)calloc(1, sizeof(mcpet_compton_table));
  T->owned = 1;
  if (!T->rayleigh || !T->compton) { mcpet_free_tables(T); return fail(MCPET_E_NOMEM, err, errlen, "load_material: out of memory"); }
  if (names[0][0] == '\0' || names[0][0] == '\n') {
    mcpet_free_tables(T);
    return fail(MCPET_E_FORMAT, err, errlen, "load_material: the first material file must be given");
  }
  char line[MCPET_NAME_LEN + 2];
  double delta_e = -99999.0;
  int N = 0;        // file points
  int NT = 0;       // file points + extrapolated points
  std::vector<double> raw;
  for (int mat = 0; mat < MCPET_MAX_MATERIALS; mat++) {
    if (names[mat][0] == '\0' || names[mat][0] == '\n') continue;
    gzFile g = gzopen(names[mat], "rb");
    if (!g) { mcpet_free_tables(T); return fail(MCPET_E_IO, err, errlen, "load_material: file %d '%s' does not exist", mat, names[mat]); }
    struct Closer { gzFile g; ~Closer() { gzclose(g); } } closer{g};
    LineReader rd(g);
#define BAD(code, ...) do { mcpet_free_tables(T); return fail(code, err, errlen, __VA_ARGS__); } while (0)
    if (!find_section(rd, "[NOMINAL DENSITY", line)) BAD(MCPET_E_FORMAT, "load_material: '%s' has no '[NOMINAL DENSITY'", names[mat]);
    if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "# %f", &T->density_nominal[mat]) != 1)
      BAD(MCPET_E_FORMAT, "load_material: bad nominal density in '%s'", names[mat]);
    if (!(density_max[mat] > 0)) {
      if (mat == 0) density_max[mat] = 0.01f * T->density_nominal[mat];   // C:1617-1618
      else continue;                                                        // C:1619-1620
    }
    int nv = 0;
    rd.gets(line, MCPET_NAME_LEN);
    if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "# %d", &nv) != 1 || nv < 2)
      BAD(MCPET_E_FORMAT, "load_material: bad number of data values in '%s'", names[mat]);
    if (mat == 0) {
      if (nv > MCPET_MAX_ENERGYBINS_RAYLEIGH)
        BAD(MCPET_E_RANGE, "load_material: too many energy bins (%d > %d)", nv, MCPET_MAX_ENERGYBINS_RAYLEIGH);
      N = nv;
    } else if (nv != N) {
      BAD(MCPET_E_FORMAT, "load_material: incorrect number of energy values in '%s': %d, expected %d", names[mat], nv, N);
    }
    rd.gets(line, MCPET_NAME_LEN);
    rd.gets(line, MCPET_NAME_LEN);
    // the file's rows (raw doubles)
    raw.assign(6 * (size_t)std::max(N, NT), 0.0);
    double e_last = -1.0;
    for (int i = 0; i < N; i++) {
      double* r = &raw[6 * (size_t)i];
      if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "  %le  %le  %le  %le  %le  %le", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5]) != 6)
        BAD(MCPET_E_FORMAT, "load_material: bad mean-free-path row %d in '%s'", i, names[mat]);
      const double d_e = r[0];
      if (i == 0 && mat == 0) T->e0 = (float)d_e;
      if (i == 0) {
        if (fabs(d_e - (double)T->e0) > 1.0e-9)
          BAD(MCPET_E_FORMAT, "load_material: incorrect first energy in '%s': %f, expected %f", names[mat], d_e, (double)T->e0);
      } else if (i == 1) {
        delta_e = d_e - e_last;
      } else if (fabs((d_e - e_last) - delta_e) / delta_e > 0.001) {
        BAD(MCPET_E_FORMAT, "load_material: the energy step is not constant in '%s' (row %d)", names[mat], i);
      }
      e_last = d_e;
    }
    if (mat == 0) {   // the grid: the files' points, continued up to e_max
      NT = N;
      if (e_max > e_last) {
        const double extra = std::ceil((e_max - e_last) / delta_e);
        if ((double)N + extra + 1.0 > (double)MCPET_MAX_ENERGYBINS_RAYLEIGH)
          BAD(MCPET_E_RANGE, "load_material: %.1f eV needs more than %d energy points", e_max, MCPET_MAX_ENERGYBINS_RAYLEIGH);
        NT = N + (int)extra + 1;   // one more point: Rayleigh sampling reads the next bin's pmax
      }
      T->num_values = NT;
      T->num_file_values = NT > N ? N : 0;
      T->woodcock = (float*)calloc(2 * (size_t)NT, sizeof(float));
      T->mfp_a = (float*)calloc(3 * (size_t)NT * MCPET_MAX_MATERIALS, sizeof(float));
      T->mfp_b = (float*)calloc(3 * (size_t)NT * MCPET_MAX_MATERIALS, sizeof(float));
      if (!T->woodcock || !T->mfp_a || !T->mfp_b) BAD(MCPET_E_NOMEM, "load_material: out of memory");
      for (int i = 0; i < NT; i++) T->woodcock[2 * i] = 99999999.99f;
      raw.resize(6 * (size_t)NT);
    }
    if (NT > N) extrapolate_rows(raw, N, NT, (double)T->e0, delta_e);
    float* A = T->mfp_a;
    float* B = T->mfp_b;
    const double dn = (double)T->density_nominal[mat];
    for (int i = 0; i < NT; i++) {
      const double* r = &raw[6 * (size_t)i];
      const double d_ray = r[1], d_com = r[2], d_tot = r[4], d_pmax = r[5];
      // C:1682-1684: minimum total mfp over materials at their maximum density.
      float temp_mfp = (float)(d_tot * dn / (double)density_max[mat]);
      if (temp_mfp < T->woodcock[2 * i]) T->woodcock[2 * i] = temp_mfp;
      // C:1688-1690: inverse mfp per unit nominal density.
      const size_t bin = (size_t)i * MCPET_MAX_MATERIALS + mat;
      A[3 * bin + 0] = (float)(1.0 / (d_tot * dn));
      A[3 * bin + 1] = (float)(1.0 / (d_com * dn));
      A[3 * bin + 2] = (float)(1.0 / (d_ray * dn));
      T->rayleigh->pmax[bin] = (float)d_pmax;
    }
    T->ide = (float)(1.0 / delta_e);   // C:1725 (1.0f/delta_e evaluated in double)
    // C:1728-1736: slopes; the last bin repeats the previous slope.
    for (int i = 0; i < NT - 1; i++) {
      const size_t bin = (size_t)i * MCPET_MAX_MATERIALS + mat;
      const size_t nxt = bin + MCPET_MAX_MATERIALS;
      for (int c = 0; c < 3; c++) B[3 * bin + c] = (float)((double)(A[3 * nxt + c] - A[3 * bin + c]) / delta_e);
    }
    {
      const size_t last = (size_t)(NT - 1) * MCPET_MAX_MATERIALS + mat, prev = last - MCPET_MAX_MATERIALS;
      for (int c = 0; c < 3; c++) B[3 * last + c] = B[3 * prev + c];
    }
    // C:1739-1746: re-base so that value = a + E*b.
    for (int i = 0; i < NT; i++) {
      const double e_i = (double)T->e0 + i * delta_e;
      const size_t bin = (size_t)i * MCPET_MAX_MATERIALS + mat;
      for (int c = 0; c < 3; c++) A[3 * bin + c] = (float)((double)A[3 * bin + c] - e_i * (double)B[3 * bin + c]);
    }
    // Rayleigh RITA data (C:1748-1782).
    for (;;) {
      rd.gets(line, MCPET_NAME_LEN);
      if (rd.eof()) BAD(MCPET_E_FORMAT, "load_material: Rayleigh data not found in '%s'", names[mat]);
      if (strstr(line, "[DATA VALUES")) break;
    }
    int nray = 0;
    if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "# %d", &nray) != 1 || nray != MCPET_NP_RAYLEIGH)
      BAD(MCPET_E_FORMAT, "load_material: the number of Rayleigh values must be %d in '%s'", MCPET_NP_RAYLEIGH, names[mat]);
    rd.gets(line, MCPET_NAME_LEN);
    for (int i = 0; i < nray; i++) {
      const int bin = MCPET_NP_RAYLEIGH * mat + i;
      int itl = 0, itu = 0;
      if (!rd.gets(line, MCPET_NAME_LEN) ||
          sscanf(line, "  %e  %e  %e  %e  %d  %d", &T->rayleigh->xco[bin], &T->rayleigh->pco[bin], &T->rayleigh->aco[bin],
                 &T->rayleigh->bco[bin], &itl, &itu) != 6)
        BAD(MCPET_E_FORMAT, "load_material: bad Rayleigh row %d in '%s'", i, names[mat]);
      // RITA interval bounds: the sampler reads pco[itl-1 .. itu-1] (+1), so itl in [1,127],
      // itu in [itl+1,128] keeps every access inside this material's 128 points.
      if (itl < 1 || itl > MCPET_NP_RAYLEIGH - 1 || itu <= itl || itu > MCPET_NP_RAYLEIGH)
        BAD(MCPET_E_RANGE, "load_material: Rayleigh interval index out of range in '%s'", names[mat]);
      T->rayleigh->itlco[bin] = (uint8_t)itl;
      T->rayleigh->ituco[bin] = (uint8_t)itu;
    }
    // Compton shells (C:1785-1814).
    for (;;) {
      rd.gets(line, MCPET_NAME_LEN);
      if (rd.eof()) BAD(MCPET_E_FORMAT, "load_material: Compton data not found in '%s'", names[mat]);
      if (strstr(line, "[NUMBER OF SHELLS")) break;
    }
    int nsh = 0;
    if (!rd.gets(line, MCPET_NAME_LEN) || sscanf(line, "# %d", &nsh) != 1 || nsh < 1 || nsh > MCPET_MAX_SHELLS)
      BAD(MCPET_E_FORMAT, "load_material: bad number of Compton shells in '%s' (max %d)", names[mat], MCPET_MAX_SHELLS);
    T->compton->noscco[mat] = nsh;
    rd.gets(line, MCPET_NAME_LEN);
    for (int i = 0; i < nsh; i++) {
      const int bin = mat + i * MCPET_MAX_MATERIALS;
      int kz, ks;
      if (!rd.gets(line, MCPET_NAME_LEN) ||
          sscanf(line, " %e  %e  %e  %d  %d", &T->compton->fco[bin], &T->compton->uico[bin], &T->compton->fj0[bin], &kz, &ks) != 5)
        BAD(MCPET_E_FORMAT, "load_material: bad Compton shell row %d in '%s'", i, names[mat]);
    }
#undef BAD
  }
  // Woodcock slope and re-base (C:1821-1829). The slope of the last point is not set by the
  // reference (heap memory); it is 0 here, which leaves the last point's value unchanged.
  for (int i = 0; i < NT - 1; i++)
    T->woodcock[2 * i + 1] = (float)((double)(T->woodcock[2 * (i + 1)] - T->woodcock[2 * i]) / delta_e);
  T->woodcock[2 * (NT - 1) + 1] = 0.0f;
  for (int i = 0; i < NT; i++)
    T->woodcock[2 * i] = (float)((double)T->woodcock[2 * i] - ((double)T->e0 + i * delta_e) * (double)T->woodcock[2 * i + 1]);
  return MCPET_OK;
}

// ------------------------------------------------------------------------------- outputs

static std::string join(const char* dir, const char* name)
{
  if (!dir || !dir[0]) return name;
  std::string s(dir);
  if (s.back() != '/') s += '/';
  return s + name;
}

static mcpet_status write_gz(const std::string& path, const void* data, size_t bytes, char* err, size_t errlen)
{
  gzFile g = gzopen(path.c_str(), "wb");
  if (!g) return fail(MCPET_E_IO, err, errlen, "cannot open '%s' for writing", path.c_str());
  const char* p = (const char*)data;
  size_t left = bytes;
  while (left) {
    unsigned chunk = left > (1u << 30) ? (1u << 30) : (unsigned)left;
    if (gzwrite(g, p, chunk) != (int)chunk) { gzclose(g); return fail(MCPET_E_IO, err, errlen, "write error on '%s'", path.c_str()); }
    p += chunk;
    left -= chunk;
  }
  gzclose(g);
  return MCPET_OK;
}

// report_PSF (MCGPU-PET.cu:3193-3322): ASCII header + first 5000 records, full binary .raw.
static mcpet_status write_psf(const std::string& path, const mcpet_config* cfg, const mcpet_detector* det,
                              const mcpet_tallies* t, double time_elapsed, char* err, size_t errlen)
{
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return fail(MCPET_E_IO, err, errlen, "report_PSF: file %s can not be opened", path.c_str());
  std::string rawp = path + ".raw";
  FILE* fb = fopen(rawp.c_str(), "w");
  if (!fb) { fclose(f); return fail(MCPET_E_IO, err, errlen, "report_PSF: file %s can not be opened", rawp.c_str()); }
  const int64_t n = t->psf_count;
  fprintf(f, "# \n");
  fprintf(f, "#     *****************************************************************************\n");
  fprintf(f, "#     ***   MCGPU-PET-compatible phase space file (MI355X photon-transport library)  ***\n");
  fprintf(f, "#     *****************************************************************************\n");
  fprintf(f, "# \n");
  fprintf(f, "#  *** SIMULATION IN THE GPU USING HIP (gfx950) ***\n");
  fprintf(f, "#\n");
  fprintf(f, "#  Phase space file (PSF) of particles arriving at the ideal cylindrical detector surrounding the object.\n");
  fprintf(f, "# \n");
  fprintf(f, "#        Number of detections in the PSF = %lld\n", (long long)n);
  fprintf(f, "#        Allocated PSF elements = %d\n", det->psf_size);
  fprintf(f, "#        PET acquistion time    = %lf s\n", cfg->acquisition_time_ps * 1.0e-12);
  fprintf(f, "#        mean life isotope      = %f\n", cfg->mean_life);
  fprintf(f, "#        Material activities:\n");
  for (int i = 0; i < MCPET_MAX_MATERIALS; i++)
    if (cfg->activity[i] > 0.0f) fprintf(f, "#                Mat %d: %f Bq\n", i + 1, cfg->activity[i]);
  fprintf(f, "#        PSF detector center    = (%.5f,%.5f,%.5f) cm\n", det->center[0], det->center[1], det->center[2]);
  fprintf(f, "#        PSF detector height and radius = %.5f, %.5f cm\n", det->height, det->radius);
  fprintf(f, "#        Input voxel file       = %s\n", cfg->voxel_file);
  fprintf(f, "# \n");
  if (det->tally_type == 1) fprintf(f, "#        PSF includes only True coincidences (Scatter not reported).\n");
  else if (det->tally_type == 2) fprintf(f, "#        PSF includes only Scatter coincidences (Trues not reported).\n");
  else fprintf(f, "#        PSF includes both True and Scatter coincidences.\n");
  if (det->tally_psf_sinogram == 1) fprintf(f, "#        Reported only PSF (Sinogram not reported).\n");
  else if (det->tally_psf_sinogram == 2) fprintf(f, "#        Reported only Sinogram (PSF not reported).\n");
  else fprintf(f, "#        Reported both PSF and Sinogram.\n");
  fprintf(f, "# \n");
  fprintf(f, "#        Reporting only the 5000 first elements below in ASCII. Complete PSF available in binary form in the file %s\n", rawp.c_str());
  fprintf(f, "#        Binary size for each reported variable:\n");
  fprintf(f, "#               - emission_time (ps): unsigned long long int \n");
  fprintf(f, "#               - travel_time (ps): float \n");
  fprintf(f, "#               - emission voxel: int \n");
  fprintf(f, "#               - energy (eV), z (cm), phi (rad), vx, vy, vz: float \n");
  fprintf(f, "#               - index1 (Flag for scatter: =0 for non-scattered, =1 for Compton, =2 for Rayleigh, and =3 for multiple scatter), index2: short int  \n");
  fprintf(f, "# \n");
  fprintf(f, "#  [emission_time (ps)]    [travel_time (ps)]    [emission voxel]  [energy (eV)]  [z (cm)]  [phi (rad)]  [vx]  [vy]  [vz]  [index1]  [index2]\n");
  fprintf(f, "# ========================================================================================================================================\n");
  const mcpet_psf_record* P = t->psf;
  for (int64_t i = 0; i < (n < 5000 ? n : 5000); i++)
    fprintf(f, "%llu %f %d %f %f %f %f %f %f %d %d\n", (unsigned long long)P[i].emission_time_ps, P[i].travel_time_ps,
            P[i].emission_absvox, P[i].energy, P[i].z, P[i].phi, P[i].vx, P[i].vy, P[i].vz, P[i].index1, P[i].index2);
  fprintf(f, "\n#        (...)\n");
  if (n > 0 && fwrite(P, sizeof(mcpet_psf_record), (size_t)n, fb) != (size_t)n) {
    fclose(f); fclose(fb);
    return fail(MCPET_E_IO, err, errlen, "report_PSF: write error on %s", rawp.c_str());
  }
  fprintf(f, "# \n");
  fprintf(f, "#   *** Simulation REPORT: ***\n");
  fprintf(f, "#       Simulated x rays:    %llu\n", (unsigned long long)t->total_histories);
  fprintf(f, "#       Simulation time [s]: %.2f\n", time_elapsed);
  if (time_elapsed > 0.000001) fprintf(f, "#       Speed [x-rays/s]:    %.2f\n\n", (double)t->total_histories / time_elapsed);
  fclose(f);
  fclose(fb);
  return MCPET_OK;
}

// One ROI voxel of report_voxels_dose (C:2505-2566): the voxel's dose and standard deviation
// (eV/g per history) as the reference computes them, with voxels below 0.1 g/cm3 reporting 0.
struct RoiVoxel {
  int64_t geometry;    // voxel index in the phantom
  int mat;             // material slot, 0-based
  double mass;         // g
  uint64_t e1, e2;     // tallied sum E and sum E^2 (eV), zeroed for low-density voxels
  double dose, sd;
};

template <class F>
static void for_each_roi_voxel(const mcpet_config* cfg, const mcpet_phantom* ph, const uint64_t* voxel_edep,
                               uint64_t total_histories, F&& f)
{
  const int DX = cfg->roi_x_max - cfg->roi_x_min + 1, DY = cfg->roi_y_max - cfg->roi_y_min + 1,
            DZ = cfg->roi_z_max - cfg->roi_z_min + 1;
  const double inv_N = 1.0 / (double)(total_histories ? total_histories : 1);
  const double voxel_volume = 1.0 / (((double)ph->inv_voxel_size[0]) * ((double)ph->inv_voxel_size[1]) * ((double)ph->inv_voxel_size[2]));
  int64_t voxel = 0;
  for (int k = 0; k < DZ; k++)
    for (int j = 0; j < DY; j++)
      for (int i = 0; i < DX; i++, voxel++) {
        RoiVoxel v;
        v.geometry = (int64_t)(i + cfg->roi_x_min) + (int64_t)(j + cfg->roi_y_min) * ph->nx +
                     (int64_t)(k + cfg->roi_z_min) * ph->nx * ph->ny;
        const float dens = ph->voxel_mat_dens[3 * v.geometry + 1];
        v.mat = (int)ph->voxel_mat_dens[3 * v.geometry] - 1;
        v.mass = dens * voxel_volume;
        const double inv_mass = 1.0 / v.mass;
        v.e1 = voxel_edep[2 * voxel];
        v.e2 = voxel_edep[2 * voxel + 1];
        const uint64_t raw1 = v.e1, raw2 = v.e2;
        if (dens < 0.1f) { v.e1 = 0; v.e2 = 0; }   // C:2523-2527: no dose reported in low-density voxels
        double dose = (double)v.e1 * inv_N;
        double sd = ((double)v.e2 * inv_N - dose * dose) * inv_N * inv_mass;
        if (sd > 0.0) sd = std::sqrt(sd);
        v.dose = dose * inv_mass;
        v.sd = sd;
        f(i, j, k, v, raw1, raw2);
      }
}

extern "C" mcpet_status mcpet_voxel_dose_report(const mcpet_config* cfg, const mcpet_phantom* ph, const uint64_t* voxel_edep,
                                                uint64_t total_histories, mcpet_dose_report* out)
{
  if (!cfg || !ph || !ph->voxel_mat_dens || !voxel_edep || !out) return MCPET_E_ARG;
  if (cfg->roi_x_min < 0 || cfg->roi_y_min < 0 || cfg->roi_z_min < 0 || cfg->roi_x_max >= ph->nx || cfg->roi_y_max >= ph->ny ||
      cfg->roi_z_max >= ph->nz || cfg->roi_x_min > cfg->roi_x_max || cfg->roi_y_min > cfg->roi_y_max || cfg->roi_z_min > cfg->roi_z_max)
    return MCPET_E_RANGE;
  std::memset(out, 0, sizeof *out);
  double edep2[MCPET_MAX_MATERIALS] = {0};
  for (int m = 0; m < MCPET_MAX_MATERIALS; m++) {   // C:2488-2495
    out->peak_dose[m] = -1.0;
    out->peak_std_dev[m] = 1.0e-15;
  }
  uint64_t total = 0;
  for_each_roi_voxel(cfg, ph, voxel_edep, total_histories, [&](int i, int j, int k, const RoiVoxel& v, uint64_t raw1, uint64_t raw2) {
    if (v.mat < 0 || v.mat >= MCPET_MAX_MATERIALS) return;
    out->mass[v.mat] += v.mass;                 // C:2512-2515: every ROI voxel of the slot, before the exclusion
    out->edep[v.mat] += (double)raw1;
    edep2[v.mat] += (double)raw2;
    out->voxels[v.mat]++;
    total += v.e1;
    if (v.dose > out->peak_dose[v.mat]) {        // C:2541-2554
      out->peak_dose[v.mat] = v.dose;
      out->peak_std_dev[v.mat] = v.sd;
      out->peak_voxel[v.mat][0] = i + cfg->roi_x_min;
      out->peak_voxel[v.mat][1] = j + cfg->roi_y_min;
      out->peak_voxel[v.mat][2] = k + cfg->roi_z_min;
      if (v.dose > out->max_dose) {
        out->max_dose = v.dose;
        out->max_dose_std_dev = v.sd;
      }
    }
  });
  const double inv_N = 1.0 / (double)(total_histories ? total_histories : 1);
  out->total_energy_kev = 0.001 * (double)total * inv_N;
  for (int m = 0; m < MCPET_MAX_MATERIALS; m++) {   // C:2614-2636
    if (!out->voxels[m]) continue;
    const double E = out->edep[m] * inv_N;
    double sd = (edep2[m] * inv_N - E * E) * inv_N;
    if (sd > 0.0) sd = std::sqrt(sd);
    out->edep[m] = E;
    out->dose[m] = E / out->mass[m];
    out->dose_std_dev[m] = sd / out->mass[m];
  }
  return MCPET_OK;
}

// report_voxels_dose (MCGPU-PET.cu:2389-2650): ASCII mid-plane, float dose .raw and 2-sigma % error .raw.
static mcpet_status write_voxel_dose(const std::string& path, const mcpet_config* cfg, const mcpet_phantom* ph,
                                     const mcpet_tallies* t, double time_elapsed, char* err, size_t errlen)
{
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return fail(MCPET_E_IO, err, errlen, "report_voxels_dose: file %s can not be opened", path.c_str());
  FILE* fm = fopen((path + ".raw").c_str(), "w");
  FILE* fs = fopen((path + "_PercentRelError2sigma.raw").c_str(), "w");
  if (!fm || !fs) {
    fclose(f); if (fm) fclose(fm); if (fs) fclose(fs);
    return fail(MCPET_E_IO, err, errlen, "report_voxels_dose: binary file %s can not be opened", path.c_str());
  }
  const int DX = cfg->roi_x_max - cfg->roi_x_min + 1, DY = cfg->roi_y_max - cfg->roi_y_min + 1,
            DZ = cfg->roi_z_max - cfg->roi_z_min + 1;
  const int z_plane = (cfg->roi_z_max + cfg->roi_z_min) / 2, z_plane_roi = z_plane - cfg->roi_z_min;
  fprintf(f, "# \n#  3D voxel dose tally (KERMA approximation: photons only, energy deposited locally)\n");
  fprintf(f, "#  *** SIMULATION IN THE GPU USING HIP (gfx950) ***\n#\n");
  fprintf(f, "#  The 3D dose deposition is reported in binary form in the .raw files (data given as 32-bit floats). \n");
  fprintf(f, "#  The %% relative error in the voxel dose at 2 standard deviations [=100*2*sigma/voxel_dose] is reported in the *_PercentRelError2sigma.raw file (32-bit floats). \n");
  fprintf(f, "#  This text output reports only the 2D dose at the Z plane z_coord = %d (z_coord in ROI = %d)\n#\n", z_plane, z_plane_roi);
  fprintf(f, "#  The total dose deposited in each material inside the ROI is reported to the standard output.\n#\n");
  fprintf(f, "#  Voxel size:  %lf x %lf x %lf = %lf cm^3\n", 1.0 / (double)ph->inv_voxel_size[0], 1.0 / (double)ph->inv_voxel_size[1],
          1.0 / (double)ph->inv_voxel_size[2],
          1.0 / (double)(ph->inv_voxel_size[0] * ph->inv_voxel_size[1] * ph->inv_voxel_size[2]));
  fprintf(f, "#  Number of voxels in the reported region of interest (ROI) X, Y and Z:\n#      %d  %d  %d\n", DX, DY, DZ);
  fprintf(f, "#  Coordinates of the ROI inside the voxel volume = X[%d,%d], Y[%d,%d], Z[%d,%d]\n#\n", cfg->roi_x_min + 1,
          cfg->roi_x_max + 1, cfg->roi_y_min + 1, cfg->roi_y_max + 1, cfg->roi_z_min + 1, cfg->roi_z_max + 1);
  fprintf(f, "#  Voxel dose units: eV/g per history\n");
  fprintf(f, "#  X rows given first, then Y, then Z. One blank line separates the different Y, and two blanks the Z values (GNUPLOT format).\n# \n");
  fprintf(f, "#    [DOSE]   [2*standard_deviation]\n# =====================================\n");
  uint64_t total_energy = 0;
  bool ok = true;
  for_each_roi_voxel(cfg, ph, t->voxel_edep, t->total_histories, [&](int i, int j, int k, const RoiVoxel& v, uint64_t, uint64_t) {
    total_energy += v.e1;
    if (k == z_plane_roi) fprintf(f, "%.6lf %.6lf\n", v.dose, 2.0 * v.sd);
    const float df = (float)v.dose;
    const float rel = v.dose > 0.0 ? 200.0f * (float)(v.sd / v.dose) : 0.0f;
    ok = ok && fwrite(&df, sizeof(float), 1, fm) == 1 && fwrite(&rel, sizeof(float), 1, fs) == 1;
    if (k == z_plane_roi && i == DX - 1) {
      fprintf(f, "\n");                       // Y rows apart
      if (j == DY - 1) fprintf(f, "\n");      // then the plane
    }
  });
  const double inv_N = 1.0 / (double)(t->total_histories ? t->total_histories : 1);
  // C:2586-2595 (one projection: a PET acquisition)
  fprintf(f, "#   ****** DOSE REPORT: TOTAL SIMULATION PERFORMANCE FOR ALL PROJECTIONS ******\n");
  fprintf(f, "#       Total number of simulated x rays: %llu\n", (unsigned long long)t->total_histories);
  fprintf(f, "#       Simulated x rays per projection:  %llu\n", (unsigned long long)t->total_histories);
  fprintf(f, "#       Total simulation time [s]:  %.2f\n", time_elapsed);
  if (time_elapsed > 0.000001) fprintf(f, "#       Total speed [x-rays/s]:  %.2f\n", (double)t->total_histories / time_elapsed);
  fprintf(f, "\n#       Total energy absorved inside the dose ROI: %.5lf keV/hist\n\n", 0.001 * (double)total_energy * inv_N);
  fclose(f);
  ok = (fclose(fm) == 0) && ok;
  ok = (fclose(fs) == 0) && ok;
  if (!ok) return fail(MCPET_E_IO, err, errlen, "report_voxels_dose: write error on %s.raw", path.c_str());
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_write_outputs(const char* dir, const mcpet_config* cfg, const mcpet_phantom* ph,
                                            const mcpet_tables* tab, const mcpet_tallies* t, double time_elapsed,
                                            char* err, size_t errlen)
{
  (void)tab;
  if (!cfg || !ph || !t) return fail(MCPET_E_ARG, err, errlen, "mcpet_write_outputs: null argument");
  const int64_t nvox = (int64_t)ph->nx * ph->ny * ph->nz;
  mcpet_status st;
  const int mode = cfg->detector.tally_psf_sinogram;
  if (mode == 0 || mode == 2) {
    if (!t->trues || !t->scatter || !t->image_trues || !t->image_scatter)
      return fail(MCPET_E_ARG, err, errlen, "mcpet_write_outputs: sinogram/image buffers missing");
    const bool bg = cfg->bg_mode != 0;   // C:544-608
    if ((st = write_gz(join(dir, bg ? "sinogram_BG_Trues.raw.gz" : "sinogram_Trues.raw.gz"), t->trues, 4 * (size_t)cfg->nbins, err, errlen))) return st;
    if ((st = write_gz(join(dir, bg ? "sinogram_BG_Scatter.raw.gz" : "sinogram_Scatter.raw.gz"), t->scatter, 4 * (size_t)cfg->nbins, err, errlen))) return st;
    if ((st = write_gz(join(dir, bg ? "BG_image_True.raw.gz" : "image_Trues.raw.gz"), t->image_trues, 4 * (size_t)nvox, err, errlen))) return st;
    if ((st = write_gz(join(dir, bg ? "BG_image_Scatter.raw.gz" : "image_Scatter.raw.gz"), t->image_scatter, 4 * (size_t)nvox, err, errlen))) return st;
  }
  if (t->spectrum) {   // C:612-616 (always written)
    std::string p = join(dir, "Energy_Sinogram_Spectrum.dat");
    FILE* f = fopen(p.c_str(), "w");
    if (!f) return fail(MCPET_E_IO, err, errlen, "cannot open '%s'", p.c_str());
    fprintf(f, "# HISTOGRAM OF ENERGIES (keV) OF DETECTED COINCIDENCE PHOTONS (before energy window selection)\n");
    fprintf(f, "#     Energy bin [keV]     Counts\n");
    for (int i = 0; i < cfg->ne; i++) fprintf(f, "%10d  %10d \n", i, t->spectrum[i]);
    fclose(f);
  }
  if ((mode == 0 || mode == 1) && t->psf) {
    mcpet_detector det = cfg->detector;
    if (det.radius < 0.0f) {   // C:316-322
      det.center[0] = ph->size_bbox[0] * 0.5f;
      det.center[1] = ph->size_bbox[1] * 0.5f;
      det.center[2] = ph->size_bbox[2] * 0.5f;
      det.radius = -det.radius;
    }
    if ((st = write_psf(join(dir, cfg->psf_file), cfg, &det, t, time_elapsed, err, errlen))) return st;
  }
  if (cfg->voxel_dose && cfg->roi_x_max > -1 && t->voxel_edep) {
    if ((st = write_voxel_dose(join(dir, cfg->dose_file), cfg, ph, t, time_elapsed, err, errlen))) return st;
  }
  return MCPET_OK;
}
