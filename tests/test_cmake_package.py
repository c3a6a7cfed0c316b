"""find_package(Tiled-MM) for CMake consumers, as the reference exports it (Tiled-MM
CMakeLists.txt:44-63, cmake/Tiled-MMConfig.cmake.in: package Tiled-MM, imported target
Tiled-MM::Tiled-MM, version 2.3.1 with SameMajorVersion compatibility).

A small consumer project (the reference's own API: make_context, gemm, get_full_device_buffer_c,
and the C ABI) is configured and linked against the in-tree build and against a
`make install PREFIX=...` tree. Nothing is run: this is the build-system contract, checked on the
CPU; the GPU tests run the same calls.
"""
import os
import shutil
import subprocess

import pytest

import tiledmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tiled-mm_amd")

CONSUMER_CMAKE = """cmake_minimum_required(VERSION 3.17)
project(consumer LANGUAGES CXX)
find_package(Tiled-MM {version} REQUIRED)
add_executable(consumer main.cpp)
target_link_libraries(consumer PRIVATE Tiled-MM::Tiled-MM)
"""

CONSUMER_MAIN = """#include <Tiled-MM/tiled_mm.hpp>
#include <Tiled-MM/util.hpp>
#include <tiledmm.h>
#include <cstdio>
int main() {
  auto ctx = gpu::make_context<double>(2, 512, 512, 512);
  double a = 1, b = 1, c = 0;
  gpu::gemm(*ctx, 'N', 'N', 1, 1, 1, 1.0, &a, 1, &b, 1, 0.0, &c, 1, false, true);
  std::printf("%zu %s\\n", ctx->get_full_device_buffer_c().size(), tmm_last_error(nullptr));
  return 0;
}
"""

pytestmark = pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")


def _configure_and_build(tmp_path, prefix, version="2.3"):
    src = tmp_path / "src"
    src.mkdir(exist_ok=True)
    (src / "CMakeLists.txt").write_text(CONSUMER_CMAKE.format(version=version))
    (src / "main.cpp").write_text(CONSUMER_MAIN)
    build = tmp_path / f"build-{version}"
    cfg = subprocess.run(["cmake", "-S", str(src), "-B", str(build), f"-DCMAKE_PREFIX_PATH={prefix}",
                          "-DCMAKE_CXX_COMPILER=g++"], capture_output=True, text=True, timeout=300)
    if cfg.returncode != 0:
        return cfg, None
    bld = subprocess.run(["cmake", "--build", str(build), "-j", "4"], capture_output=True, text=True, timeout=300)
    return cfg, bld


def _check_linked(build_dir):
    exe = build_dir / "consumer"
    assert exe.exists()
    out = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert "libtiledmm.so" in out, out


def test_find_package_in_tree(tmp_path):
    assert os.path.exists(os.path.join(tiledmm.LIB_DIR, "cmake", "Tiled-MM", "Tiled-MMConfig.cmake")), \
        "build the library first (make -C tiled-mm_amd)"
    cfg, bld = _configure_and_build(tmp_path, PKG)  # PKG/lib/cmake/Tiled-MM
    assert cfg.returncode == 0, cfg.stdout + cfg.stderr
    assert bld.returncode == 0, bld.stdout + bld.stderr
    _check_linked(tmp_path / "build-2.3")


def test_find_package_installed_tree(tmp_path):
    prefix = tmp_path / "prefix"
    r = subprocess.run(["make", "-C", PKG, "install", f"PREFIX={prefix}"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    for rel in ("lib/libtiledmm.so", "lib/cmake/Tiled-MM/Tiled-MMConfig.cmake",
                "lib/cmake/Tiled-MM/Tiled-MMTargets.cmake", "include/tiledmm.h", "include/Tiled-MM/tiled_mm.hpp",
                "include/Tiled-MM/mm_handle.hpp"):
        assert (prefix / rel).exists(), rel
    cfg, bld = _configure_and_build(tmp_path, prefix)
    assert cfg.returncode == 0, cfg.stdout + cfg.stderr
    assert bld.returncode == 0, bld.stdout + bld.stderr
    _check_linked(tmp_path / "build-2.3")


def test_find_package_rejects_other_major_version(tmp_path):
    cfg, _ = _configure_and_build(tmp_path, PKG, version="3.0")
    assert cfg.returncode != 0
    assert "Tiled-MM" in cfg.stderr
