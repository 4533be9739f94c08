// cli.cpp — reference-compatible command line: `mcgpu_pet MCGPU-PET.in [anything -> BG mode]`.
//
// Mirrors main() of the reference (MCGPU-PET.cu:77-840): read the .in file, load the
// phantom and material tables, run the transport on the GPU, write the same output
// files and print the same summary numbers. Paths in the .in file are relative to the
// current directory, as in the reference. Exit codes follow the reference: -1 for
// missing input/device, -2 for malformed inputs, -3 for output errors.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mcpet.h"

static int exit_code(mcpet_status st)
{
  switch (st) {
    case MCPET_OK: return 0;
    case MCPET_E_IO: return -1;
    case MCPET_E_HIP: return -1;
    default: return -2;
  }
}

// report_materials_dose (MCGPU-PET.cu:2664-2725), printed to stdout.
static void report_materials_dose(const mcpet_config& cfg, const mcpet_phantom& ph, const mcpet_tables& tab,
                                  const uint64_t* mat_edep, uint64_t total_histories)
{
  double vol = 1.0 / (((double)ph.inv_voxel_size[0]) * ((double)ph.inv_voxel_size[1]) * ((double)ph.inv_voxel_size[2]));
  double mass[MCPET_MAX_MATERIALS] = {0};
  const long long nv = (long long)ph.nx * ph.ny * ph.nz;
  for (long long v = 0; v < nv; v++) mass[(int)ph.voxel_mat_dens[3 * v] - 1] += (double)ph.voxel_mat_dens[3 * v + 1] * vol;
  printf("\n\n          *** MATERIALS TOTAL DOSE TALLY REPORT ***\n\n");
  printf("              Total number of simulated x rays: %llu\n", (unsigned long long)total_histories);
  printf("\t [MAT]  [DOSE eV/g/hist]  [2*std_dev]  [Rel_error 2*std_dev, %%]  [E_dep eV/hist]  [DOSE mGy]  [Material mass g]  [Material file name]\n");
  const double inv_N = 1.0 / (double)(total_histories ? total_histories : 1);
  for (int i = 0; i < MCPET_MAX_MATERIALS; i++) {
    if (tab.density_nominal[i] < 0.0f) break;
    const char* name = strrchr(cfg.material_files[i], '/');
    name = name ? name + 1 : cfg.material_files[i];
    double Edep = (double)mat_edep[2 * i] * inv_N;
    double sd = sqrt(((double)mat_edep[2 * i + 1] * inv_N - Edep * Edep) * inv_N);
    double rel = Edep > 0.0 ? sd / Edep : 0.0;
    double m = mass[i] > 0.00001 ? mass[i] : 0.00001;
    printf("\t%d\t%.5lf\t\t%.5lf\t\t%.2lf\t\t%.5lf\t\t%.5lf\t\t%.5lf\t\t%s\n", i + 1, Edep / m, 2.0 * sd / m, 200.0 * rel, Edep,
           (double)mat_edep[2 * i] / m * (1.0e3 / 6.2415e15), mass[i], name);
  }
}

// The standard-output part of report_voxels_dose (MCGPU-PET.cu:2597-2641): energy in the ROI,
// peak voxel dose and the per-material ROI dose table.
static void report_voxels_dose(const mcpet_config& cfg, const mcpet_phantom& ph, const uint64_t* voxel_edep,
                               uint64_t total_histories)
{
  mcpet_dose_report r;
  if (mcpet_voxel_dose_report(&cfg, &ph, voxel_edep, total_histories, &r) != MCPET_OK) return;
  printf("\n\n          *** VOXEL ROI DOSE TALLY REPORT ***\n\n");
  printf("\n              Total energy absorved inside the dose deposition ROI: %.5lf keV/hist\n", r.total_energy_kev);
  printf("              Maximum voxel dose (+-2 sigma): %lf +- %lf eV/g per history.\n", r.max_dose, r.max_dose_std_dev);
  printf("              Dose deposited in the different materials inside the input ROI computed post-processing the 3D voxel dose results:\n\n");
  printf("  [MAT]  [DOSE_ROI eV/g/hist]  [2*std_dev]  [Rel error %%]  [Peak voxel dose]  [2*std_dev]  [Rel error %%]  [Peak voxel coord]  [E_dep eV/hist]  [MASS_ROI g]  [NUM_VOXELS_ROI]\n");
  printf(" ===============================================================================================================================================================================\n");
  for (int i = 0; i < MCPET_MAX_MATERIALS; i++) {
    if (!r.voxels[i]) continue;
    double rel = 0.0, rel_peak = 0.0;
    if (r.dose[i] > 0.0) {
      rel = r.dose_std_dev[i] / r.dose[i];
      rel_peak = r.peak_std_dev[i] / r.peak_dose[i];
    }
    printf("\t%d\t%.5lf\t%.5lf\t%.3lf\t\t%.5lf\t%.5lf\t%.3lf\t(%d,%d,%d)\t\t%.5lf\t%.5lf\t%u\n", i + 1, r.dose[i],
           2.0 * r.dose_std_dev[i], 200.0 * rel, r.peak_dose[i], 2.0 * r.peak_std_dev[i], 200.0 * rel_peak, r.peak_voxel[i][0],
           r.peak_voxel[i][1], r.peak_voxel[i][2], r.edep[i], r.mass[i], r.voxels[i]);
  }
  printf("\n");
}

int main(int argc, char** argv)
{
  if (argc < 2) {
    printf("\n\n   !!read_input ERROR!! Input file name not given as an execution parameter!! Try again...\n\n");
    return -1;
  }
  char err[1024];
  mcpet_config cfg;
  mcpet_status st = mcpet_read_input(argv[1], &cfg, err, sizeof err);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", err); return exit_code(st); }
  cfg.bg_mode = (argc == 3) ? 1 : 0;   // C:81-84
  printf("\n    MCGPU-PET-compatible photon transport on MI355X (%s)\n", mcpet_version());
  printf("    -- Seed used in the execution '%d'\n", cfg.seed);
  printf("                   PET acquistion time = %lf s\n", cfg.acquisition_time_ps * 1.0e-12);
  printf("                     mean life isotope = %f\n", cfg.mean_life);
  mcpet_phantom ph;
  st = mcpet_load_voxels(cfg.voxel_file, &cfg, &ph, err, sizeof err);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", err); return exit_code(st); }
  printf("       Number of voxels in the input geometry file: %d x %d x %d =  %d\n", ph.nx, ph.ny, ph.nz, ph.nx * ph.ny * ph.nz);
  mcpet_tables tab;
  st = mcpet_load_materials(cfg.material_files, ph.density_max, &tab, err, sizeof err);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", err); return exit_code(st); }
  mcpet_ctx* ctx = nullptr;
  st = mcpet_create(&cfg, &ph, &tab, &ctx);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", mcpet_last_error(nullptr)); return exit_code(st); }
  mcpet_detector det;
  mcpet_get_detector(ctx, &det);
  printf("\n\n    ** INPUT PET DETECTOR PARAMETERS  **\n");
  printf("           Phase Space File (PSF) size =  %d\n", det.psf_size);
  printf("                   PSF detector center = (%.5f,%.5f,%.5f) cm\n", det.center[0], det.center[1], det.center[2]);
  printf("        PSF detector height and radius = %.5f, %.5f cm\n", det.height, det.radius);

  auto t0 = std::chrono::steady_clock::now();
  mcpet_run_params rp{};
  rp.zero_tallies = 1;
  st = mcpet_run(ctx, &rp);
  if (!st) st = mcpet_sync(ctx);
  auto t1 = std::chrono::steady_clock::now();
  if (st) { printf("\n\n !!Kernel execution failed while simulating particle tracks!! %s\n\n", mcpet_last_error(ctx)); return -1; }
  double elapsed = std::chrono::duration<double>(t1 - t0).count();
  printf("\n\n    -- MONTE CARLO LOOP finished: time tallied in main program: %.3f s\n", elapsed);

  const long long nvox = (long long)ph.nx * ph.ny * ph.nz;
  std::vector<int32_t> trues(cfg.nbins), scatter(cfg.nbins), imt(nvox), imsc(nvox), spec(cfg.ne > 0 ? cfg.ne : 1);
  std::vector<mcpet_psf_record> psf;
  const int mode = cfg.detector.tally_psf_sinogram;
  if ((mode == 0 || mode == 1) && cfg.detector.psf_size > 0) psf.resize(cfg.detector.psf_size);
  long long n_roi = 0;
  if (cfg.voxel_dose && cfg.roi_x_max > -1)
    n_roi = (long long)(cfg.roi_x_max - cfg.roi_x_min + 1) * (cfg.roi_y_max - cfg.roi_y_min + 1) * (cfg.roi_z_max - cfg.roi_z_min + 1);
  std::vector<uint64_t> vdose(2 * (n_roi > 0 ? n_roi : 1)), mdose(2 * MCPET_MAX_MATERIALS);
  mcpet_tallies t{};
  t.trues = trues.data();
  t.scatter = scatter.data();
  t.image_trues = imt.data();
  t.image_scatter = imsc.data();
  t.spectrum = spec.data();
  t.psf = psf.empty() ? nullptr : psf.data();
  t.psf_capacity = (int64_t)psf.size();
  t.voxel_edep = n_roi > 0 ? vdose.data() : nullptr;
  t.material_edep = mdose.data();
  st = mcpet_fetch(ctx, &t);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", mcpet_last_error(ctx)); return exit_code(st); }
  printf("                                  Total number of histories simulated = %llu (note: each annihilation produces 2 histories)\n",
         (unsigned long long)t.total_histories);
  if (t.psf_requested <= cfg.detector.psf_size)
    printf("                                  Number of coincidences stored in the PSF = %lld\n\n\n", (long long)t.psf_requested / 2);
  else
    printf("\n                       WARNING:  %lld photons arrived at the detector but only %d elements could be stored in the PSF!!\n",
           (long long)t.psf_requested, cfg.detector.psf_size);
  st = mcpet_write_outputs(nullptr, &cfg, &ph, &tab, &t, elapsed, err, sizeof err);
  if (st) { printf("\n\n   !!ERROR!! %s\n\n", err); return -3; }
  if (n_roi > 0) report_voxels_dose(cfg, ph, vdose.data(), t.total_histories);
  if (cfg.material_dose) report_materials_dose(cfg, ph, tab, mdose.data(), t.total_histories);
  long long simt = 0, simsc = 0;
  for (long long v = 0; v < nvox; v++) { simt += imt[v]; simsc += imsc[v]; }
  printf("\n\n       ****** MCGPU-PET SINOGRAM INFORMATION ******\n\n");
  printf("           >>> Number of radial bins = %i\n", cfg.nrad);
  printf("           >>> Number of angular bins = %i\n", cfg.nangles);
  printf("           >>> Total number of sinograms = %i\n", cfg.nsinos);
  printf("           >>> Maximum Ring Difference = %i\n", cfg.mrd);
  printf("           >>> SPAN = %i\n", cfg.span);
  printf("           >>> Number segments = %i\n", cfg.nseg);
  printf("           >>> Number z slices = %i\n\n", cfg.nzs);
  printf("           >>> Number of true coincidences in the sinogram = %lld  \n", (long long)t.sum_trues);
  printf("           >>> Number of scatter coincidences in the sinogram = %lld  \n", (long long)t.sum_scatter);
  printf("           >>> Size of the 3D sinogram = %i X %i X %i\n", cfg.nrad, cfg.nangles, cfg.nsinos);
  printf("\n\n       ****** MCGPU-PET IMAGES INFORMATION ******\n\n");
  printf("           >>> Number of voxels = (%i,%i,%i)\n", ph.nx, ph.ny, ph.nz);
  printf("           >>> Number of true counts in the image of trues = %lld  \n", simt);
  printf("           >>> Number of scatter counts in the image of scatter = %lld  \n", simsc);
  printf("\n\n       ****** TOTAL SIMULATION PERFORMANCE ******\n\n");
  printf("          >>> Time spent in the Monte Carlo transport only: %.3f s.\n", elapsed);
  printf("          >>> Total number of simulated histories (x-rays):  %llu  (note: each positron annihilation produces 2 histories)\n",
         (unsigned long long)t.total_histories);
  if (elapsed > 1e-7) {
    printf("          >>> Total speed Monte Carlo transport only [x-rays/s]:  %.2f\n", (double)t.total_histories / elapsed);
    printf("          >>> Detected coincidences (trues + scatter) per second:  %.2f\n\n", (double)(t.sum_trues + t.sum_scatter) / elapsed);
  }
  mcpet_destroy(ctx);
  mcpet_free_tables(&tab);
  mcpet_free_phantom(&ph);
  return 0;
}
