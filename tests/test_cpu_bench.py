"""The benchmark's host-side helpers: kernel-name shortening, the rocprofv3
summary, and the one-GPU contract (any other --gpus exits with a message)."""
import json
import os
import subprocess
import sys

import bench

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def test_short_kernel_names():
    assert bench._short_kernel("void gmvs::pm_candidates<6, 2, 0, false, 2>(gmvs::PMProblem const*, "
                               "int)") == "gmvs::pm_candidates<6, 2, 0, false, 2>"
    assert bench._short_kernel('"gmvs::(anonymous namespace)::fuse_check<float>(gmvs::(anonymous '
                               'namespace)::FuseView const*, int)"') == "gmvs::fuse_check<float>"
    assert bench._short_kernel("gmvs::pm_select(gmvs::PMProblem const*)") == "gmvs::pm_select"


def test_summarize_profile(tmp_path):
    stats = tmp_path / "p" / "x_kernel_stats.csv"
    stats.parent.mkdir()
    rows = ["Name,Calls,TotalDurationNs,AverageNs,Percentage,MinNs,MaxNs"]
    # 22 torch kernels ahead of one gmvs kernel: the gmvs kernel is still listed
    for i in range(22):
        rows.append(f'"at::native::k{i}(int)",10,{(100 - i) * 1000000},{(100 - i) * 100000},1,1,1')
    rows.append('"void gmvs::(anonymous namespace)::fuse_fill(int)",4,2000,500,0.1,400,600')
    stats.write_text("\n".join(rows) + "\n")
    pmc = tmp_path / "p" / "x_counter_collection.csv"
    pmc.write_text("Kernel_Name,Counter_Name,Counter_Value,Dispatch_Id,Grid_Size\n"
                   '"gmvs::pm_select(int)",FETCH_SIZE,10,1,256\n'
                   '"gmvs::pm_select(int)",FETCH_SIZE,30,2,256\n'
                   '"__amd_rocclr_copyBuffer",FETCH_SIZE,524308,3,131072\n')
    out = bench.summarize_profile(str(tmp_path / "p"))
    lines = out.split("\n")
    assert any(l.startswith("gmvs::fuse_fill ") and " 4 " in l for l in lines)
    assert not any(l.startswith("at::native::k21") for l in lines)  # 21st and later torch rows
    sel = [l for l in lines if l.startswith("gmvs::pm_select")]
    assert sel and sel[0].split()[1] == "2" and float(sel[0].split()[2]) == 20.0  # mean per dispatch
    assert any("524308.0 FETCH_SIZE __amd_rocclr_copyBuffer" in l for l in lines)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
