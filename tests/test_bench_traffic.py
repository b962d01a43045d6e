"""bench.py's HBM-traffic post-processing (--traffic-from), on the CPU: bytes
from the request-size counters of two rocprofv3 --pmc passes, per kernel the
median of its last three dispatches, no calibration factor."""
import csv
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402

HDR = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
R2G32 = "void oa3::k3_r2g<float, 10, 10, 10, 10, 256, true, false, oa3::Lay<5> >(oa::FftCtx)"
R2G32_RHO = "void oa3::k3_r2g<float, 10, 10, 10, 10, 256, true, true, oa3::Lay<5> >(oa::FftCtx)"
ZB64 = "void oa2::k2_zbwd<double, 10, 10, true>(oa::FftCtx)"
COPY = "__amd_rocclr_copyBuffer"


def write(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HDR)
        w.writerows(rows)


def test_traffic_from_request_size_counters(tmp_path):
    kernels = [COPY, R2G32, R2G32_RHO, ZB64]
    rd, wr = [], []
    for d in range(16):  # four dispatches of each kernel, interleaved
        k = kernels[d % 4]
        n32, n64, n128 = 3 + d, 5, 100 + 10 * d   # reads of 32, 64, 128 bytes
        rd += [(d, k, "TCC_EA0_RDREQ_sum", n32 + n64 + n128), (d, k, "TCC_EA0_RDREQ_32B_sum", n32),
               (d, k, "TCC_EA0_RDREQ_64B_sum", n64), (d, k, "TCC_EA0_RDREQ_128B_sum", n128)]
        w64, wall = 40 + d, 50 + d                  # 64-byte writes among all writes
        wr += [(d, k, "TCC_EA0_WRREQ_sum", wall), (d, k, "TCC_EA0_WRREQ_64B_sum", w64)]
    write(str(tmp_path / "rd"), rd)
    write(str(tmp_path / "wr"), wr)
    out = tmp_path / "t.json"

    def prec(name):
        return "mixed" if "<float" in name else ("fp64" if "<double" in name else None)

    t = bench.traffic_from_counters([str(tmp_path / "rd"), str(tmp_path / "wr")], str(out), prec)
    assert json.load(open(out)) == json.loads(json.dumps(t))

    def bytes_of(d):  # what dispatch d moved: 32 n32 + 64 n64 + 128 n128 read, 64 w64 + 32 (rest)
        return 32 * (3 + d) + 64 * 5 + 128 * (100 + 10 * d), 64 * (40 + d) + 32 * 10

    # R2G32 ran as dispatches 1, 5, 9, 13: the median of the last three is dispatch 9
    r, w = bytes_of(9)
    assert t["detail"][R2G32]["read_bytes"] == r and t["detail"][R2G32]["write_bytes"] == w
    assert t["mixed"]["k_plane_r2g"] == r + w
    r, w = bytes_of(10)  # the RHO variant is a stage of its own
    assert t["mixed"]["k_plane_r2g_rho"] == r + w
    r, w = bytes_of(11)
    assert t["fp64"]["k_zbwd"] == r + w
    assert t["detail"][ZB64]["reads_of_other_sizes"] == 0
    # the copy check reports the copy dispatch with the most read requests (dispatch 12)
    r, w = bytes_of(12)
    assert t["copy_check"]["read_bytes"] == r and t["copy_check"]["write_bytes"] == w
