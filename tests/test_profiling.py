"""Host logic of arcanefem_amd.profiling: kernel-name shortening and the
copy-calibrated HBM traffic computed from rocprofv3 FETCH_SIZE / WRITE_SIZE
CSV files (synthetic CSVs with the columns rocprofv3 writes)."""

import csv

import pytest

from arcanefem_amd import profiling as P

FIELDS = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, grid, value) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Grid_Size": grid, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": value})


def test_short_name():
    assert P.short_name("void afem::k_spmv_sell<double, double, 0>(int, int const*, double*)") == \
        "k_spmv_sell<double, double, 0>"
    assert P.short_name("afem::k_pair_plan(int const*, int)") == "k_pair_plan"
    assert P.short_name("void afem::k_x<4, (afem::Mode)1>(afem::Parts)") == "k_x<4, (Mode)1>"


def test_pmc_summary_scaling(tmp_path):
    copy = P.COPY_KERNEL
    kern = "void afem::k_assemble_laplace_lplan<4, 4>(double const*, int)"
    f = tmp_path / "f.csv"
    w = tmp_path / "w.csv"
    # the 256 MiB copy reads half its bytes on FETCH_SIZE and all on WRITE_SIZE
    write_csv(f, "FETCH_SIZE", [(copy, 512, 1.5), (copy, 131072, 131072.0), (kern, 2146816, 300000.0),
                                (kern, 2146816, 310000.0), (kern, 2146816, 305000.0)])
    write_csv(w, "WRITE_SIZE", [(copy, 512, 0.03), (copy, 131072, 262144.0), (kern, 2146816, 250000.0),
                                (kern, 2146816, 250000.0), (kern, 2146816, 250000.0)])
    s = P.pmc_summary(str(f), str(w), 256 << 20)
    assert s["fetch_factor"] == pytest.approx(2.0)
    assert s["write_factor"] == pytest.approx(1.0)
    k = s["kernels"]["k_assemble_laplace_lplan<4, 4>"]
    assert k["dispatches"] == 3
    assert k["fetch_bytes"] == pytest.approx(305000.0 * 2.0 * 1024)  # median, scaled
    assert k["traffic_bytes"] == pytest.approx((305000.0 * 2.0 + 250000.0) * 1024)
    md = P.pmc_markdown(s, "cmd", "tet128_seed0")
    assert "k_assemble_laplace_lplan<4, 4>" in md and "2.0000" in md


def test_pmc_summary_needs_calibration(tmp_path):
    f = tmp_path / "f.csv"
    w = tmp_path / "w.csv"
    write_csv(f, "FETCH_SIZE", [("void afem::k(int)", 64, 1.0)])
    write_csv(w, "WRITE_SIZE", [("void afem::k(int)", 64, 1.0)])
    with pytest.raises(SystemExit):
        P.pmc_summary(str(f), str(w), 256 << 20)


def test_counter_table(tmp_path):
    kern = "void afem::k_assemble_laplace_brick<4, 16>(double const*, int)"
    other = "void afem::k_spmv_sell16<double, double, 0, 4, 256>(int)"
    a = tmp_path / "a.csv"
    b = tmp_path / "b.csv"
    write_csv(a, "SQ_INSTS_VALU", [(kern, 64, 10.0), (kern, 64, 30.0), (kern, 64, 20.0), (other, 64, 5.0)])
    write_csv(b, "SQ_LDS_BANK_CONFLICT", [(kern, 64, 7.0), (P.COPY_KERNEL, 64, 1.0)])
    t = P.counter_table([str(a), str(b)], match="assemble")
    assert set(t) == {"k_assemble_laplace_brick<4, 16>"}
    assert t["k_assemble_laplace_brick<4, 16>"]["SQ_INSTS_VALU"] == (20.0, 3)
    assert t["k_assemble_laplace_brick<4, 16>"]["SQ_LDS_BANK_CONFLICT"] == (7.0, 1)
    md = P.counters_markdown(t, "cmd")
    assert "SQ_INSTS_VALU" in md and "| 3 |" in md
