"""The reference's file contract end to end on the GPU: the command-line program
(`mcgpu-pet_amd/bin/mcgpu_pet file.in [x]`, the replacement of main(), MCGPU-PET.cu:77-840)
run on the reference's sample input (MCGPU-PET.in + phantom_9x9x9cm.vox, with a fixed seed, a
short acquisition, a PSF that fits and the voxel dose switched on) writes, for the same seed,
exactly the tallies the library returns: the int32 sinograms and emission images (.raw.gz),
the energy spectrum text, the PSF text header and .raw records, and the voxel dose .raw and
2-sigma %-error .raw files; its printed summaries match the fetched totals. A second argument
selects the background mode and its BG_ file names (C:81-84, C:580-594)."""
import gzip
import os
import re
import subprocess

import numpy as np
import pytest

import mcpet
from conftest import ROOT, SAMPLE
from harness import psf_sorted

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "mcgpu-pet_amd", "bin", "mcgpu_pet")


def sample_input(tmp_path, seed, acq_s, psf_size, roi):
    """The sample .in with the given seed, acquisition, PSF capacity, voxel dose ROI (1-based,
    inclusive) and absolute input paths."""
    txt = open(os.path.join(SAMPLE, "MCGPU-PET.in")).read()
    lines = txt.splitlines()
    out = []
    for l in lines:
        if "# RANDOM SEED" in l:
            l = f"{seed}    # RANDOM SEED"
        elif "# TOTAL PET SCAN ACQUISITION TIME" in l:
            l = f"{acq_s}    # TOTAL PET SCAN ACQUISITION TIME  [seconds]"
        elif "# PHASE SPACE FILE SIZE" in l:
            l = f" {psf_size}    # PHASE SPACE FILE SIZE (MAXIMUM NUMBER OF ELEMENTS)"
        elif "# TALLY 3D VOXEL DOSE?" in l:
            l = "YES    # TALLY 3D VOXEL DOSE? [YES/NO]"
        elif "# VOXEL DOSE ROI: X-index" in l:
            l = f"  {roi[0]}  {roi[1]}    # VOXEL DOSE ROI: X-index min max (first voxel has index 1)"
        elif "# VOXEL DOSE ROI: Y-index" in l:
            l = f"  {roi[2]}  {roi[3]}    # VOXEL DOSE ROI: Y-index min max"
        elif "# VOXEL DOSE ROI: Z-index" in l:
            l = f"  {roi[4]}  {roi[5]}    # VOXEL DOSE ROI: Z-index min max"
        elif l.strip().startswith("phantom_9x9x9cm.vox"):
            l = os.path.join(SAMPLE, "phantom_9x9x9cm.vox") + "    # VOXEL GEOMETRY FILE"
        elif l.strip().startswith("./materials/"):
            name, rest = l.split("#", 1)
            l = os.path.join(SAMPLE, name.strip()[2:]) + "    #" + rest
        out.append(l)
    p = tmp_path / "MCGPU-PET.in"
    p.write_text("\n".join(out) + "\n")
    return str(p)


def library_run(path, bg=False):
    cfg = mcpet.read_input(path)
    cfg.bg_mode = 1 if bg else 0
    ph = mcpet.load_voxels(cfg.voxel_file.decode(), cfg)
    tab = mcpet.load_materials(cfg, ph.density_max())
    with mcpet.Simulation(cfg, ph, tab) as sim:
        sim.run()
        out = sim.fetch()
    return cfg, ph, tab, out


def gz_int32(p):
    return np.frombuffer(gzip.open(p).read(), np.int32)


def test_cli_sample_input_files_equal_the_library_results(tmp_path):
    run_dir = tmp_path / "cli"
    run_dir.mkdir()
    path = sample_input(tmp_path, seed=8642, acq_s=1.0, psf_size=2_000_000, roi=(2, 8, 1, 9, 3, 7))
    print("CLI: sample input, seed 8642, 1 s, PSF 2e6, voxel dose ROI x2-8 y1-9 z3-7", flush=True)
    r = subprocess.run([EXE, path], cwd=run_dir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    cfg, ph, tab, lib = library_run(path)
    # sinograms and emission images
    for name, key in (("sinogram_Trues.raw.gz", "trues"), ("sinogram_Scatter.raw.gz", "scatter"),
                      ("image_Trues.raw.gz", "image_trues"), ("image_Scatter.raw.gz", "image_scatter")):
        assert np.array_equal(gz_int32(run_dir / name), lib[key]), name
    assert lib["trues"].sum() > 1000
    # spectrum
    rows = [l.split() for l in open(run_dir / "Energy_Sinogram_Spectrum.dat") if not l.startswith("#")]
    assert np.array_equal(np.array([int(c) for _, c in rows]), lib["spectrum"])
    assert [int(b) for b, _ in rows] == list(range(cfg.ne))
    # PSF: the text header's count and the binary records (a multiset: the slot order follows
    # the device's atomic order)
    raw = np.fromfile(run_dir / "MCGPU_PET.psf.raw", mcpet.PSF_DTYPE)
    assert len(raw) == len(lib["psf"]) > 1000
    assert np.array_equal(psf_sorted(raw), psf_sorted(lib["psf"]))
    head = open(run_dir / "MCGPU_PET.psf").read()
    assert f"Number of detections in the PSF = {len(raw)}" in head
    # voxel dose: the library's writer on the fetched tallies writes the same bytes
    ref_dir = tmp_path / "lib"
    ref_dir.mkdir()
    mcpet.write_outputs(str(ref_dir), cfg, ph, tab, lib, 1.0)
    for name in ("mc-gpu_dose.dat.raw", "mc-gpu_dose.dat_PercentRelError2sigma.raw"):
        a, b = (run_dir / name).read_bytes(), (ref_dir / name).read_bytes()
        assert len(a) == 4 * 7 * 9 * 5 and a == b, name
    assert np.fromfile(run_dir / "mc-gpu_dose.dat.raw", np.float32).max() > 0
    rep = mcpet.voxel_dose_report(cfg, ph, lib["voxel_edep"], lib["total_histories"])
    # printed summaries
    out = r.stdout
    assert f"Total number of histories simulated = {lib['total_histories']}" in out
    assert f"Number of true coincidences in the sinogram = {lib['trues'].sum()}" in out
    assert f"Number of scatter coincidences in the sinogram = {lib['scatter'].sum()}" in out
    assert f"Number of coincidences stored in the PSF = {len(raw) // 2}" in out
    m = re.search(r"Total energy absorved inside the dose deposition ROI: ([0-9.]+) keV/hist", out)
    assert m and float(m.group(1)) == float(f"{rep['total_energy_kev']:.5f}")
    assert "MATERIALS TOTAL DOSE TALLY REPORT" in out and "Peak voxel dose" in out


def test_cli_background_mode_writes_the_bg_files(tmp_path):
    run_dir = tmp_path / "cli"
    run_dir.mkdir()
    path = sample_input(tmp_path, seed=97, acq_s=0.5, psf_size=1_000_000, roi=(1, 9, 1, 9, 1, 9))
    print("CLI: sample input in background mode (second argument), seed 97, 0.5 s", flush=True)
    r = subprocess.run([EXE, path, "BG"], cwd=run_dir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert not (run_dir / "sinogram_Trues.raw.gz").exists()
    _, _, _, lib = library_run(path, bg=True)
    for name, key in (("sinogram_BG_Trues.raw.gz", "trues"), ("sinogram_BG_Scatter.raw.gz", "scatter"),
                      ("BG_image_True.raw.gz", "image_trues"), ("BG_image_Scatter.raw.gz", "image_scatter")):
        assert np.array_equal(gz_int32(run_dir / name), lib[key]), name
    assert lib["trues"].sum() + lib["scatter"].sum() > 100
