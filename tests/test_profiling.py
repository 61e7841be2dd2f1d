"""CPU tests of the rocprofv3 summary tool: kernel naming (the two lattice
instances stay apart) and the traffic scaling by the copy kernel."""
import csv

import pytest

from pscbfm_amd import profiling as pr

BIT = "void pscbfm::(anonymous namespace)::k_check<1, 2, 2>(pscbfm::SysDesc, unsigned int)"
U8 = "void pscbfm::(anonymous namespace)::k_check<0, 2, 2>(pscbfm::SysDesc, unsigned int)"
P2U8 = "void pscbfm::(anonymous namespace)::k_perform<0>(pscbfm::SysDesc, unsigned int, unsigned int)"
BATCH = ("void pscbfm::(anonymous namespace)::k_check_batch<1, 2, 4>(pscbfm::SysDesc const*, pscbfm::Tile const*, "
         "HIP_vector_type<unsigned int, 2u>*, unsigned int*, unsigned int)")
COPY = "pscbfm::(anonymous namespace)::k_copy_u4(HIP_vector_type<unsigned int, 4u> const*, HIP_vector_type<unsigned int, 4u>*, unsigned long)"


def test_kernel_names_keep_lattice_instances_apart():
    assert pr._short(BIT) == "k_check" and pr._short(U8) == "k_check.u8"
    assert pr._short(P2U8) == "k_perform.u8" and pr._short(BATCH) == "k_check_batch"
    assert pr._short(COPY) == "k_copy_u4"
    assert pr.label("k_check.u8").endswith("uint8 lattice") and pr.label("k_check") == "pass 1 (k_check)"


def _counter_dir(tmp_path, name, rows):
    d = tmp_path / name
    d.mkdir()
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (k, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})
    return d


def test_traffic_is_scaled_by_the_copy_kernel(tmp_path):
    copy_bytes = 1 << 20
    # the copy kernel reads and writes 1 MiB; its counters report half / all of it
    fetch = _counter_dir(tmp_path, "f", [(COPY, "FETCH_SIZE", 512.0), (BIT, "FETCH_SIZE", 100.0),
                                         (BIT, "FETCH_SIZE", 300.0), (U8, "FETCH_SIZE", 1000.0)])
    write = _counter_dir(tmp_path, "w", [(COPY, "WRITE_SIZE", 1024.0), (BIT, "WRITE_SIZE", 50.0),
                                         (BIT, "WRITE_SIZE", 50.0), (U8, "WRITE_SIZE", 10.0)])
    t = pr.traffic(fetch, write, copy_bytes)
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    bit = t["kernels"]["k_check"]
    assert bit["read_bytes"] == pytest.approx(200.0 * 2.0 * 1024) and bit["write_bytes"] == pytest.approx(50.0 * 1024)
    u8 = t["kernels"]["k_check.u8"]
    assert u8["traffic_bytes"] == pytest.approx(1000.0 * 2.0 * 1024 + 10.0 * 1024)
    # steady state: the mean over the last dispatch only
    last = pr.traffic(fetch, write, copy_bytes, last=1)
    assert last["kernels"]["k_check"]["read_bytes"] == pytest.approx(300.0 * 2.0 * 1024)
