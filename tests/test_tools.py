"""Host tools: kernel-name shortening of the rocprof summaries kept under
profiles/, and bench.py's single-GPU contract."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_kernel_names():
    from glass_amd.rocprof_summary import _short
    cases = {
        "void glass::eval_kernel<512, 256, false, 1>(int const*, int const*, glass::Items, glass::BatchDev, "
        "int const*, double*, double*)": "eval_kernel<512,256,false,1>",
        "accept_staged_kernel(int, glass::Items, glass::BatchDev, HIP_vector_type<double, 2u> const*, int const*, "
        "int const*, double const*)": "accept_staged_kernel",
        "void glass::(anonymous namespace)::noise_logl_kernel(int, int, double const*)": "noise_logl_kernel",
        "glass::wdm::resp_kernel(glass::wdm::Args)": "resp_kernel",
        "__amd_rocclr_copyBuffer": "__amd_rocclr_copyBuffer",
    }
    for name, short in cases.items():
        assert _short(name) == short, name


def test_trace_summary(tmp_path):
    from glass_amd import rocprof_summary
    csv = tmp_path / "run_kernel_trace.csv"
    csv.write_text("Kernel_Name,Grid_Size_X,Grid_Size_Y,Grid_Size_Z,Workgroup_Size_X,Workgroup_Size_Y,"
                   "Workgroup_Size_Z,Start_Timestamp,End_Timestamp\n"
                   '"void glass::k<8>(int)",512,1,1,64,1,1,1000,3000\n'
                   '"void glass::k<8>(int)",512,1,1,64,1,1,5000,9000\n'
                   '"void glass::k<8>(int)",64,1,1,64,1,1,0,500\n')
    import contextlib
    import io
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rocprof_summary.trace(str(csv))
    lines = buf.getvalue().splitlines()
    assert lines[0].startswith("kernel | blocks")
    assert "k<8> | 8 | 2 | 3.00 | 0.006" in lines and "k<8> | 1 | 1 | 0.50 | 0.001" in lines


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
