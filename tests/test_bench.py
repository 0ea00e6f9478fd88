"""bench.py's contract logic on the CPU: which kernel the roofline prices, the
traffic figure it takes from the committed counter summary, and the refusal
of more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_roofline_prices_the_dominant_kernel_with_a_byte_model():
    # ms, launches per class; the scheduler holds the most time but has no
    # byte model, so the DAG kernel is priced and the scheduler is named
    kt = {"scheduler": (500.0, 4), "dag_dp": (400.0, 500), "dag_edges": (180.0, 500), "ripup": (60.0, 500)}
    kb = {"dag_dp": 4.0e10, "dag_edges": 2.4e11}
    r = bench.roofline(kt, kb, 1)
    assert r["kernel"] == "dag_dp" and r["symbol"] == "k_dag_net"
    assert r["largest_class"] == "scheduler"
    assert r["achieved"] == pytest.approx(4.0e10 / 0.4 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS)
    assert r["bytes_per_launch"] == pytest.approx(4.0e10 / 500)
    assert r["avg_launch_us"] == pytest.approx(800.0)
    assert set(r["classes"]) == set(kt)
    assert "GB/s" in r["classes"]["dag_edges"] and "GB/s" not in r["classes"]["scheduler"]


def test_roofline_prices_the_scheduler_when_it_has_bytes():
    # with the scheduler's device-counted bytes it is the priced leader; the
    # DAG kernel keeps its counter traffic in the class list
    kt = {"scheduler": (500.0, 4), "dag_dp": (400.0, 500)}
    kb = {"scheduler": 2.0e10, "dag_dp": 4.0e10}
    r = bench.roofline(kt, kb, 2)
    assert r["kernel"] == "scheduler" and r["symbol"] is None and "largest_class" not in r
    assert r["achieved"] == pytest.approx(2.0e10 / 0.5 / 1e9)
    assert r["traffic"] is None
    assert r["classes"]["dag_dp"]["traffic"] == bench._traffic("k_dag_net", 2)[0]


def test_roofline_without_largest_class_when_priced_kernel_leads():
    r = bench.roofline({"dag_dp": (400.0, 500), "scheduler": (100.0, 4)}, {"dag_dp": 4.0e10}, 1)
    assert r["kernel"] == "dag_dp" and "largest_class" not in r


@pytest.mark.parametrize("config", [1, 2])
def test_traffic_comes_from_the_committed_counter_summary(config):
    name = "traffic_config%d_batched.txt" % config
    path = os.path.join(ROOT, "profiles", name)
    want = None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == "k_dag_net":
            want = float(f[2])
    assert want and want > 0
    assert bench._traffic("k_dag_net", config) == (want, "profiles/" + name)
    assert bench._traffic("no_such_kernel", config) == (None, None)
    assert bench._traffic("k_dag_net", 7) == (None, None)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def _csv(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(",".join(header) + "\n")
        for r in rows:
            f.write(",".join('"%s"' % v for v in r) + "\n")


def test_kernel_names_from_demangled_symbols():
    assert bench._short("void igr_dev::(anonymous namespace)::k_dag_net<2>(igr_dev::DagArgs, int)") == "k_dag_net"
    assert bench._short("void igr_dev::(anonymous namespace)::k_r_claim<16>(BArgs, long)", True) == "k_r_claim<16>"
    assert bench._short("__amd_rocclr_copyBuffer(char*, char*)") == "__amd_rocclr_copyBuffer"


def test_counter_traffic_is_scaled_by_the_copy_probe(tmp_path, capsys):
    # The copy probe reads and writes 1 GiB; FETCH_SIZE reads half of it
    # (factor 2), WRITE_SIZE all of it (factor 1).
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _csv(str(tmp_path / "f" / "a" / "x_counter_collection.csv"), hdr,
         [("k_copy_probe(int*)", "FETCH_SIZE", 524288), ("k_dag_net<2>(A)", "FETCH_SIZE", 900),
          ("k_dag_net<2>(A)", "FETCH_SIZE", 1100)])
    _csv(str(tmp_path / "w" / "a" / "x_counter_collection.csv"), hdr,
         [("k_copy_probe(int*)", "WRITE_SIZE", 1048576), ("k_dag_net<2>(A)", "WRITE_SIZE", 500),
          ("k_dag_net<2>(A)", "WRITE_SIZE", 500)])
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"))
    out = capsys.readouterr().out
    assert "factors 2.0000 (fetch), 1.0000 (write)" in out
    assert "traffic k_dag_net 2560000" in out            # 1000 KiB * 2 + 500 KiB * 1
    assert "traffic k_copy_probe %d" % (2 << 30) in out


def test_lds_conflict_rate(tmp_path, capsys):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _csv(str(tmp_path / "l" / "x_counter_collection.csv"), hdr,
         [("k_dag_edges_wave(A)", "SQ_LDS_BANK_CONFLICT", 10), ("k_dag_edges_wave(A)", "SQ_LDS_IDX_ACTIVE", 400),
          ("k_dag_edges_wave(A)", "SQ_LDS_BANK_CONFLICT", 10), ("k_dag_edges_wave(A)", "SQ_LDS_IDX_ACTIVE", 600),
          ("k_presum<16>(B)", "SQ_LDS_BANK_CONFLICT", 0), ("k_presum<16>(B)", "SQ_LDS_IDX_ACTIVE", 0)])
    bench.summarize_lds(str(tmp_path / "l"))
    out = capsys.readouterr().out.splitlines()
    edges = [l for l in out if l.startswith("k_dag_edges_wave")][0].split()
    assert edges[1] == "2" and edges[-1] == "0.0200"
    assert [l for l in out if l.startswith("k_presum")][0].endswith("no LDS")


def test_roofline_without_kernel_timing():
    r = bench.roofline({}, {}, 1)
    assert r["kernel"] is None and r["frac"] is None and "note" in r
