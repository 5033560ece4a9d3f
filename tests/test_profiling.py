"""rocprofv3 reduction used for profiles/ and bench.py's roofline.traffic:
kernel-name parsing, stage labels, and the copy-calibrated HBM bytes per
stage from two counter runs (synthetic counter files with known values)."""
import csv
import os

from peasoup_amd import profiling as pr


def test_short_name_keeps_template_arguments():
    assert pr.short_name("void (anonymous namespace)::k_fft_pass<1024, 16, 1, true>(psx::fft::PassArgs)") == \
        "k_fft_pass<1024, 16, 1, true>"
    assert pr.short_name("psx::batch::(anonymous namespace)::k_harm_peaks<4>(float const*, long, "
                         "psx::batch::Windows, psx::fft::TrialDesc const*, psx::batch::Cand16*, unsigned int, "
                         "unsigned int*)") == "batch::k_harm_peaks<4>"


def test_stage_labels():
    assert pr.stage_of("k_fft_pass<1024, 16, 1, true>") == "fft_pass0_resample"
    assert pr.stage_of("k_fft_pass<4096, 4, 0, false>") == "fft_pass"
    assert pr.stage_of("batch::k_spec_post<4>") == "spec_post"
    assert pr.stage_of("batch::k_spec_rows") == "spec_post"
    assert pr.stage_of("batch::k_spec_edge") is None
    assert pr.stage_of("batch::k_harm_peaks<4>") == "harm_peaks"
    assert pr.stage_of("batch::k_pool_peaks") == "pool_peaks"
    assert pr.stage_of("batch::k_harm_bound<4>") == "harm_bound"
    assert pr.stage_of("k_convert<float>") is None


def _counter_csv(path, counter, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for disp, grid, name, kib in rows:
            w.writerow([disp, grid, name, counter, kib])


def test_traffic_is_copy_calibrated_and_skips_the_prologue(tmp_path):
    gib_kib = float(1 << 20)  # 1 GiB in KiB
    # FETCH_SIZE reads half the bytes (factor 0.5), WRITE_SIZE all of them
    fetch = [(1, 1048576, "void k_convert<float>(float const*)", 0.5 * gib_kib),
             (2, 262144, "void k_fft_pass<1024, 16, 1, false>(psx::fft::PassArgs)", 7.0),  # prologue FFT
             (3, 8388608, "void k_fft_pass<1024, 16, 1, true>(psx::fft::PassArgs)", 0.5 * 3.0 * 1024),
             (4, 8388608, "void k_fft_pass<4096, 4, 0, false>(psx::fft::PassArgs)", 0.5 * 5.0 * 1024),
             (5, 4194304, "void psx::batch::k_spec_rows(float2 const*)", 0.5 * 2.0 * 1024),
             (6, 270336, "void psx::batch::k_spec_edge(float2 const*)", 1.0)]
    write = [(d, g, n, v * 2.0 if "convert" in n else v) for d, g, n, v in fetch]
    _counter_csv(str(tmp_path / "f" / "a" / "f_counter_collection.csv"), "FETCH_SIZE", fetch)
    _counter_csv(str(tmp_path / "w" / "a" / "w_counter_collection.csv"), "WRITE_SIZE", write)
    t = pr.traffic(str(tmp_path / "f"), str(tmp_path / "w"), trials=4)
    assert abs(t["calibration"]["FETCH_SIZE"]["factor"] - 0.5) < 1e-12
    assert abs(t["calibration"]["WRITE_SIZE"]["factor"] - 1.0) < 1e-12
    st = t["stages"]
    assert set(st) == {"fft_pass0_resample", "fft_pass1", "spec_post"}
    assert st["fft_pass0_resample"]["FETCH_SIZE"] == 3.0 * 1024 * 1024
    assert st["fft_pass1"]["FETCH_SIZE"] == 5.0 * 1024 * 1024
    assert st["spec_post"]["bytes_per_launch"] == 2.0 * 1024 * 1024 + 1.0 * 1024 * 1024
    assert st["spec_post"]["bytes_per_trial"] == st["spec_post"]["bytes_per_launch"] / 4


def test_issue_reports_instructions_per_wave_and_issue_share(tmp_path):
    # one prologue FFT (before the batched launches) and one batched launch
    # of each of two stages; every counter in one file
    rows = []
    for disp, grid, name, waves, valu, lds, vmem, gui in (
            (1, 262144, "void k_fft_pass<1024, 16, 1, false>(psx::fft::PassArgs)", 10, 1, 1, 1, 8.0),
            (2, 8388608, "void k_fft_pass<1024, 16, 1, true>(psx::fft::PassArgs)", 1000, 1200000, 100000, 60000,
             8.0 * 10000),
            (3, 4194304, "void psx::batch::k_spec_rows(float2 const*)", 500, 1000000, 30000, 30000, 8.0 * 5000)):
        for c, v in (("SQ_WAVES", waves), ("SQ_INSTS_VALU", valu), ("SQ_INSTS_LDS", lds),
                     ("SQ_INSTS_VMEM_RD", vmem), ("GRBM_GUI_ACTIVE", gui)):
            rows.append((disp, grid, name, c, v))
    path = tmp_path / "s" / "a" / "s_counter_collection.csv"
    os.makedirs(os.path.dirname(str(path)))
    with open(str(path), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writerows(rows)
    st = pr.issue(str(tmp_path / "s"), cus=100, xcds=8)
    assert set(st) == {"fft_pass0_resample", "spec_post"}
    p0 = st["fft_pass0_resample"]
    assert p0["cycles"] == 10000 and p0["valu_per_wave"] == 1200 and p0["lds_per_wave"] == 100
    assert abs(p0["valu_issue"] - 1200000 / (100 * 10000)) < 1e-12
    assert abs(st["spec_post"]["lds_issue"] - 30000 / (100 * 5000)) < 1e-12
