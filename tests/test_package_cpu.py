"""CMake package: `make install` lays out headers, library and package
files, and a downstream project builds against librett::librett with
find_package(librett) (the reference's consumer interface,
cmake/librett-config.cmake.in), linking the mangled librett.h API."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CONSUMER_CMAKE = """cmake_minimum_required(VERSION 3.16)
project(consumer CXX)
find_package(librett 0.1 REQUIRED)
add_executable(consumer main.cpp)
target_link_libraries(consumer librett::librett)
"""

CONSUMER_MAIN = """#include <librett.h>
#include <librett/librett_c.h>
int main() {
  librettHandle h;
  int dim[2] = {4, 4}, perm[2] = {1, 0};
  librett_gpuStream_t s = 0;
  librettResult r = librettPlan(&h, 2, dim, perm, 8, s);
  (void)r;
  librettDestroy(h);
  return librett_c_version() ? 0 : 1;
}
"""


@pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not available")
def test_install_and_find_package(tmp_path):
    prefix = tmp_path / "prefix"
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "librett_amd"), "install", "PREFIX=%s" % prefix],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for rel in ("include/librett/librett.h", "include/librett/librett_c.h", "lib/librett.so",
                "lib/cmake/librett/librett-config.cmake", "lib/cmake/librett/librett-config-version.cmake"):
        assert (prefix / rel).exists(), rel
    src = tmp_path / "consumer"
    src.mkdir()
    (src / "CMakeLists.txt").write_text(CONSUMER_CMAKE)
    (src / "main.cpp").write_text(CONSUMER_MAIN)
    build = tmp_path / "build"
    r = subprocess.run(["cmake", "-S", str(src), "-B", str(build), "-DCMAKE_PREFIX_PATH=%s" % prefix,
                        "-DCMAKE_CXX_COMPILER=g++"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run(["cmake", "--build", str(build)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (build / "consumer").exists()
    # A newer version than the package is refused.
    (src / "CMakeLists.txt").write_text(CONSUMER_CMAKE.replace("librett 0.1", "librett 9.0"))
    r = subprocess.run(["cmake", "-S", str(src), "-B", str(tmp_path / "build2"), "-DCMAKE_PREFIX_PATH=%s" % prefix,
                        "-DCMAKE_CXX_COMPILER=g++"], capture_output=True, text=True)
    assert r.returncode != 0
