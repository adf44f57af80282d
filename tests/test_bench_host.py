"""bench.py's host side, on the CPU: the one-GPU guard, the kernel names of
each launch shape, and the committed per-kernel counters it reports as
`roofline.traffic` (a name that drifted from the counters file would turn
that field null without an error)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_kernel_names_of_each_launch_shape():
    import bench
    big = dict(waves8=1, waves16=1, waves32=1, waves64=4, tile_waves=8, tile_chunk=85, tile_big=1)
    assert bench.kernels(big) == [("me_tile_kernel<8,big>", (8, 16, 32, 64))]
    small = dict(big, tile_big=0, tile_waves=4, waves32=2, waves64=8)
    assert bench.kernels(small) == [("me_search_kernel<64,8>", (64,)),
                                    ("me_search_kernel<32,2>", (32,)),
                                    ("me_tile_kernel<4>", (8, 16))]
    per_call = dict(small, tile_waves=0)
    names = [k for k, _ in bench.kernels(per_call)]
    assert names == ["me_search_kernel<64,8>", "me_search_kernel<32,2>",
                     "me_search_kernel<16,1>", "me_search_kernel<8,1>"]
    # every block size is searched by exactly one kernel in each shape
    for t in (big, small, per_call):
        sizes = sorted(s for _, ss in bench.kernels(t) for s in ss)
        assert sizes == [8, 16, 32, 64]


def test_committed_counters_cover_the_default_kernel():
    import bench
    t = json.load(open(os.path.join(ROOT, "profiles", "me_kernels_counters.json")))
    tiling = dict(t["tiling"])
    # the library's built-in launch shape (include/tcme.h), which bench.py reports
    assert tiling == dict(waves8=1, waves16=1, waves32=1, waves64=4, tile_waves=8,
                          tile_chunk=85, tile_big=1)
    name = bench.kernels(tiling)[0][0]
    c = bench.measured_counters(name, tiling)
    assert c is not None, f"no counters for {name}"
    # HBM bytes (copy-calibrated fetch + write) and VALU instructions per launch
    assert all(v > 0 for v in c.values() if isinstance(v, (int, float)))
    # another chunk size is another kernel shape: no counters are claimed for it
    assert bench.measured_counters(name, dict(tiling, tile_chunk=23)) is None
