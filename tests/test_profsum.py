"""The rocprofv3 summaries (kunquant_amd.profsum) on small hand-made CSVs:
timeline arithmetic (GPU-busy fraction, kernels in flight), the VALU rate
against the 614.4 G instr/s issue peak, and the copy-calibrated traffic."""
import csv
import json

from kunquant_amd import profsum


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_trace_timeline(tmp_path, capsys):
    p = tmp_path / "trace.csv"
    hdr = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]
    # two runs of 3 launches, back to back: [0,100) [50,150) [200,300) then +1000
    rows = []
    for r, base in enumerate((0, 1000)):
        rows += [[3 * r + 1, "kun_m_s0_k0(kun::StageArgs)", base + 0, base + 100],
                 [3 * r + 2, "void (anonymous namespace)::rank_rows<float, 1024, 8, 0>(kun_cs::View)", base + 50, base + 150],
                 [3 * r + 3, "kun_m_s1_k0(kun::StageArgs)", base + 200, base + 300]]
    rows.append([99, "at::native::copy_kernel", 5000, 5100])  # not the library's: not in a run
    _write(p, hdr, rows)
    profsum.main(["trace", str(p), "--launches", "3"])
    out = capsys.readouterr().out
    runs = [line for line in out.splitlines() if line.startswith("| 0 |") or line.startswith("| 1 |")]
    assert len(runs) == 2
    cells = [c.strip() for c in runs[0].split("|")[1:-1]]
    # wall 300 ns, busy 250 ns (union), idle 50 ns, summed 300 ns over 250 busy
    assert cells[1] == "3" and cells[3] == "0.833" and cells[5] == "1.20"
    assert "rank_rows<float, 1024, 8, 0>" in out


def test_valu_rate_and_json(tmp_path, capsys):
    p = tmp_path / "pmc.csv"
    hdr = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Counter_Name", "Counter_Value"]
    rows = []
    for d in (1, 2):
        for n, v in (("SQ_INSTS_VALU", 614.4e3), ("SQ_WAVES", 10), ("SQ_ACTIVE_INST_VALU", 50),
                     ("SQ_WAVE_CYCLES", 100)):
            rows.append([d, "kun_m_s0_k0(kun::StageArgs)", 0, 2000, n, v])
    _write(p, hdr, rows)
    out = tmp_path / "v.json"
    profsum.main(["valu", str(p), "--json", str(out)])
    rec = json.loads(out.read_text())
    k = rec["kernels"]["kun_m_s0_k0"]
    # 614.4e3 instructions in 2 us = 307.2 G/s = half the issue peak
    assert k["dispatches"] == 2 and abs(k["valu_frac"] - 0.5) < 1e-9
    row = [line for line in capsys.readouterr().out.splitlines() if line.startswith("| kun_m_s0_k0")][0]
    cells = [c.strip() for c in row.split("|")[1:-1]]
    # avg us, fraction of peak, VALU-active per wave-cycle
    assert cells[2] == "2.0" and cells[5] == "0.500" and cells[6] == "0.500"
    assert rec["runtime_sha256"] and len(rec["runtime_sha256"]) == 64


def test_traffic_scaled_by_copy(tmp_path):
    hdr = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Counter_Name", "Counter_Value"]
    gib_kib = (1 << 30) / 1024
    _write(tmp_path / "f.csv", hdr, [[1, "kun_cs_f32_s0_k0(x)", 0, 1, "FETCH_SIZE", 1000.0],
                                     [2, "void (anonymous namespace)::scale_sums<float, 1>(x)", 0, 1, "FETCH_SIZE", 10.0],
                                     [3, "void (anonymous namespace)::scale_div<float, 1, 1>(x)", 0, 1, "FETCH_SIZE", 20.0]])
    _write(tmp_path / "w.csv", hdr, [[1, "kun_cs_f32_s0_k0(x)", 0, 1, "WRITE_SIZE", 500.0],
                                     [3, "void (anonymous namespace)::scale_div<float, 1, 1>(x)", 0, 1, "WRITE_SIZE", 30.0]])
    # the copy reads half of what it moves on the fetch counter, all of it on the write counter
    _write(tmp_path / "cf.csv", hdr, [[1, "copyBuffer", 0, 1, "FETCH_SIZE", gib_kib / 2]])
    _write(tmp_path / "cw.csv", hdr, [[1, "copyBuffer", 0, 1, "WRITE_SIZE", gib_kib]])
    out = tmp_path / "t.json"
    profsum.main(["traffic", str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                  str(tmp_path / "cw.csv"), "--module", "cs_f32", "--stocks", "64", "--days", "8",
                  "--json", str(out)])
    rec = json.loads(out.read_text())
    k = rec["kernels"]
    assert k["kun_cs_f32_s0_k0"] == {"fetch_bytes": 1000.0 * 1024 * 2, "write_bytes": 500.0 * 1024}
    assert k["scale_sums + scale_div"] == {"fetch_bytes": 30.0 * 1024 * 2, "write_bytes": 30.0 * 1024}
    assert rec["scale_fetch"] == 2.0 and rec["scale_write"] == 1.0
