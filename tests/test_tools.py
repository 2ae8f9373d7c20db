"""Host-side tests of the profiling summarisers (tt-fort_amd/trace_summary.py
and tt-fort_amd/pmc_summary.py) on small synthetic rocprofv3 CSVs."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tt-fort_amd")


def _run(script, *args):
    r = subprocess.run([sys.executable, os.path.join(PKG, script), *args], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_trace_summary(tmp_path):
    # three dispatches: A (10 us), B (20 us) after a 30 us gap, A (10 us) back to back
    rows = [("kA", 0, 10000, 256), ("kB", 40000, 60000, 512), ("kA", 60000, 70000, 256)]
    p = tmp_path / "t_kernel_trace.csv"
    with open(p, "w") as f:
        f.write("Kernel_Name,Start_Timestamp,End_Timestamp,Grid_Size_X,Grid_Size_Y,Grid_Size_Z\n")
        for name, t0, t1, g in rows:
            f.write("%s,%d,%d,%d,1,1\n" % (name, t0, t1, g))
    out = _run("trace_summary.py", str(p), "kA")
    first = out.splitlines()[0].split()
    assert first[:2] == ["dispatches", "3"]
    assert abs(float(first[3]) - 0.07) < 1e-9      # span 70 us
    assert abs(float(first[5]) - 0.04) < 1e-9      # busy 40 us
    assert "gaps > 20us: 1 totalling 0.03 ms" in out
    # the gap is attributed to the kernels around it
    assert any(ln.startswith("kA") and "-> kB" in ln and ln.split()[-2] == "1" for ln in out.splitlines())
    # per-kernel table: kA twice at 10 us, kB once at 20 us
    ka = [ln for ln in out.splitlines() if ln.startswith("kA ") and "->" not in ln][0].split()
    assert ka[1] == "2" and abs(float(ka[3]) - 10.0) < 1e-9


def test_pmc_summary(tmp_path):
    p = tmp_path / "c_counter_collection.csv"
    with open(p, "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n")
        # 50 dispatches of kX: the first ten (warm-up) differ from the last 40
        for d in range(50):
            f.write("%d,kX,1024,FETCH_SIZE,%f\n" % (d, 100.0 if d < 10 else 8.0 + (d % 2)))
        f.write("50,kY,64,FETCH_SIZE,3.0\n")
    out = _run("pmc_summary.py", str(p)).splitlines()
    kx = [ln for ln in out if ln.startswith("kX")][0]
    assert "n=  50" in kx and "mean_last40=8.5" in kx and "min=8.0" in kx and "max=9.0" in kx
    ky = [ln for ln in out if ln.startswith("kY")][0]
    assert "n=   1" in ky and "mean_last40=3.0" in ky
