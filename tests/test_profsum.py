"""The rocprofv3 summaries that bench.py reads numbers from (CPU only)."""
import csv

from paratc_amd import profsum


def _pmc_csv(path, rows, counter):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, grid, value in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Counter_Name": counter, "Counter_Value": value})


def test_short_name_keeps_template_arguments():
    assert profsum.short_name("void ptc::k2_fft_z<512, 8, true>(double2*, int) [clone .kd]") == "k2_fft_z<512, 8, true>"
    assert profsum.short_name("(anonymous namespace)::k_copy_nt(double const*, double*, long)") == "k_copy_nt"


def test_totals_scale_by_the_copy_and_leave_it_out(tmp_path):
    # the copy of 1 MiB each way reads 512 KiB of FETCH_SIZE and 1024 KiB of WRITE_SIZE
    fetch = [("k_copy_nt(double*)", 64, 512.0), ("k_a(double*)", 64, 10.0), ("k_a(double*)", 64, 30.0),
             ("k_b<2>(double*)", 128, 5.0)]
    write = [("k_copy_nt(double*)", 64, 1024.0), ("k_a(double*)", 64, 4.0), ("k_a(double*)", 64, 4.0),
             ("k_b<2>(double*)", 128, 1.0)]
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    _pmc_csv(f, fetch, "FETCH_SIZE")
    _pmc_csv(w, write, "WRITE_SIZE")
    n, rb, wb = profsum.totals(str(f), str(w), float(1 << 20))
    assert n == 3
    assert rb == (10.0 + 30.0 + 5.0) * 1024.0 * 2.0
    assert wb == (4.0 + 4.0 + 1.0) * 1024.0 * 1.0


def test_totals_leave_out_large_launches_and_named_kernels(tmp_path):
    rows = [("k_copy_nt(double*)", 64, 512.0), ("k_a(double*)", 64, 10.0), ("k_a(double*)", 5000, 30.0),
            ("k_b<512>(double*)", 128, 5.0), ("k_c<64>(double*)", 128, 7.0)]
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    _pmc_csv(f, rows, "FETCH_SIZE")
    _pmc_csv(w, [(n, g, 2.0 * v) for n, g, v in rows], "WRITE_SIZE")
    n, rb, wb = profsum.totals(str(f), str(w), float(1 << 20), max_grid=1000, skip=("<512",))
    assert n == 2
    assert rb == (10.0 + 7.0) * 1024.0 * 2.0
    assert wb == (20.0 + 14.0) * 1024.0 * 1.0


def test_read_value_and_read_traffic(tmp_path):
    md = tmp_path / "s.md"
    md.write_text("# title\n\nstep_bytes_counted 1234.5\nprose step_bytes_counted 9 more\n\n"
                  "| k_a<1> @ 64 | 2 | 1.0 | 1.0 | 10 | 20 | 30 |\n"
                  "| k_a<1> @ 128 | 2 | 1.0 | 1.0 | 10 | 20 | 50 |\n")
    assert profsum.read_value(str(md), "step_bytes_counted") == 1234.5
    assert profsum.read_value(str(md), "missing") is None
    assert profsum.read_value(str(tmp_path / "none.md"), "step_bytes_counted") is None
    assert profsum.read_traffic(str(md), "k_a<1>") == 50.0


def test_bench_finds_the_traffic_of_every_pass_the_search_can_pick():
    """bench.py's roofline.traffic looks the dominant pass's kernel up in the
    committed 1024x256x512 traffic summary by name: every pass of the plans
    the search picks there (8- or 16-column z tiles, one-shot or persistent x
    passes, the complex-line y-solve) has its row."""
    import os
    import bench

    path = os.path.join(bench.ROOT, "profiles", "poisson_1024x256x512_hbm_traffic.md")
    for z_tile in (8, 16):
        for persist in (0, 1):
            plan = {"x_fast": True, "z_fast": True, "z_tile": z_tile, "persist": persist, "streaming": 1,
                    "tri": "partitioned", "chunk_rows": 8, "tri_complex_lines": 16, "tri_lds_lines": 0,
                    "lanes_per_block": 16, "slab_planes": 0}
            for kernel in bench.stage_kernels(plan, 1024, 512):
                t = profsum.read_traffic(path, kernel)
                assert t is not None and 2.1e9 < t < 2.7e9, (plan, kernel, t)
