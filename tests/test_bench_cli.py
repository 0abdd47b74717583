"""bench.py command line: one GPU only (no torch or GPU needed for the check)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    """--gpus above 1 exits with status 2 and says why, before touching a GPU."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_counter_files_are_read_per_kernel(tmp_path):
    """The live traffic passes read rocprofv3's counter-collection CSV: values
    per dispatch, grouped by kernel, other counters ignored."""
    sys.path.insert(0, ROOT)
    import bench

    f = tmp_path / "pmc_counter_collection.csv"
    f.write_text(
        "Correlation_Id,Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
        '1,1,"void irismpc::share_gemm_rot_kernel<4, 2, 4, 1>(irismpc::GemmArgs)",FETCH_SIZE,352000.5\n'
        '2,2,"void irismpc::share_gemm_rot_kernel<4, 2, 4, 1>(irismpc::GemmArgs)",FETCH_SIZE,176000\n'
        "3,3,__amd_rocclr_copyBuffer,FETCH_SIZE,1048610\n"
        "4,4,__amd_rocclr_copyBuffer,WRITE_SIZE,2097152\n")
    vals = bench.read_counters([str(f)], "FETCH_SIZE")
    gemm = [v for k, vs in vals.items() if "share_gemm_rot_kernel" in k for v in vs]
    assert gemm == [352000.5, 176000.0]
    assert vals["__amd_rocclr_copyBuffer"] == [1048610.0]
    # 2 GiB copied: FETCH_SIZE reads half the bytes in KiB, so its scale is 2
    assert abs(bench.PROBE_COPY_BYTES / (1048610 * 1024.0) - 2.0) < 1e-3
