"""The library's host code under AddressSanitizer, without a GPU.

Every HIP source is compiled for the host only (no device code), linked against the
stand-in runtime of tests/host_asan/mock_runtime.cpp (device memory in host memory, kernels
not run, every launch's grid, block and dynamic LDS checked against the hardware limits),
and tests/host_asan/driver.cpp runs plans, the single-bandwidth complex64 and FP64
transforms, batched correlation and peak search, and grouped transforms with their tuner,
graph capture, schedule timing and graph cache (including the bandwidths 32, 40, 5, 55, 40
whose Stockham FFT lengths the grouped GPU tests cover).  A heap error, a use of a destroyed
stream, event or graph, or an out-of-range launch fails the test.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "pysofft_amd", "csrc")
HERE = os.path.join(ROOT, "tests", "host_asan")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="needs hipcc and ROCm clang++")
def test_host_paths_under_address_sanitizer(tmp_path):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs, procs = [], []
    for f in srcs:
        o = str(tmp_path / (f[:-4] + ".o"))
        objs.append(o)
        # host-only objects: the sanitizer applies to host code only
        cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-host-only", "--offload-arch=gfx950",
               "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer", "-w",
               "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, f), "-o", o]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p, f in zip(procs, srcs):
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, (f, out.decode(errors="replace")[-2000:])
    nm = subprocess.run(["nm", "-u"] + objs, capture_output=True, text=True, check=True).stdout
    fat = sorted({ln.split()[-1] for ln in nm.splitlines() if "__hip_fatbin_" in ln})
    (tmp_path / "fatbins.inc").write_text("".join(f"char {s}[16] = {{0}};\n" for s in fat))
    shutil.copy(os.path.join(HERE, "mock_runtime.cpp"), tmp_path / "mock_runtime.cpp")
    for name, src in (("mock_runtime.o", str(tmp_path / "mock_runtime.cpp")),
                      ("driver.o", os.path.join(HERE, "driver.cpp"))):
        subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-fsanitize=address", "-fno-omit-frame-pointer",
                        "-I", os.path.join(ROOT, "include"), "-I", str(tmp_path), "-c", src,
                        "-o", str(tmp_path / name)], check=True)
    exe = str(tmp_path / "driver")
    subprocess.run([CLANG, "-fsanitize=address", "-o", exe, str(tmp_path / "driver.o"),
                    str(tmp_path / "mock_runtime.o")] + objs + ["-lpthread"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("launches ") and int(last.split()[1]) > 0, r.stdout[-2000:]
