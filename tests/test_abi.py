"""The drop-in boundary, checked without a GPU.

  * libtiledmm.so loads and exports every entry point include/tiledmm.h declares;
  * the C header compiles as C99, the C++ headers (include/Tiled-MM/) compile and a program using
    the reference's API (make_context, gemm, get_full_device_buffer_c, the four element types)
    links against libtiledmm.so, and the library exports the explicit instantiations;
  * pure host entry points answer without touching the GPU.
"""
import ctypes
import os
import re
import subprocess

import pytest

import oracle
import tiledmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tiledmm.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(tmm_\w+)\(", text, flags=re.M)))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads_and_exports_header():
    tiledmm.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    exported = exported_symbols(tiledmm.LIB_PATH)
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_cxx_instantiations_exported():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", tiledmm.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout.replace("> >", ">>")
    for t in ("float", "double", "std::complex<float>", "std::complex<double>"):
        assert f"void gpu::gemm<{t}>(gpu::mm_handle<{t}>&" in out, t
        assert f"gpu::mm_handle<{t}>::mm_handle(int, int, int, int)" in out, t
        assert f"gpu::mm_handle<{t}>::get_full_device_buffer_c()" in out, t


def test_tile_policy_entry_point_matches_oracle():
    L = tiledmm.load()
    for dim, mx in [(2048, 512), (40000, 8192), (24000, 6144), (32000, 6144), (50, 4), (5, 4), (12345, 5000)]:
        assert L.tmm_optimal_tile_size(dim, 0, mx) == oracle.optimal_tile_size(dim, 0, mx)


def test_null_handle_is_an_error():
    L = tiledmm.load()
    assert L.tmm_destroy(None) == 0
    assert L.tmm_set_num_streams(None, 2) == 1
    st = L.tmm_create(None, 1, 2, 64, 64, 64)
    assert st == 1 and b"null" in L.tmm_last_error(None)


def test_c_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "tiledmm.h"\nint main(void) { tmm_handle* h = 0; (void)h; return TMM_SUCCESS; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o",
                    str(tmp_path / "t.o")], check=True)


CXX_PROGRAM = r"""
#include <Tiled-MM/tiled_mm.hpp>
#include <Tiled-MM/util.hpp>
#include <complex>
template <typename T> void use() {
  auto ctx = gpu::make_context<T>(2, 512, 512, 512);
  auto ctx2 = gpu::make_context<T>();
  T* a = gpu::malloc_pinned<T>(16, T(1));
  gpu::gemm(*ctx, 'N', 'T', 4, 4, 4, T(1), a, 4, a, 4, T(0), a, 4, false, true);
  gpu::gemm(*ctx, 'n', 'n', 4, 4, 4, T(1), a, 4, a, 4, T(0), a, 4);
  auto& full = ctx->get_full_device_buffer_c();
  (void)full.data(); (void)full.size();
  auto tiles = ctx->optimal_tile_sizes(100, 100, 100);
  (void)std::get<0>(tiles);
  ctx->set_tile_sizes(64, 64, 64);
  ctx->set_num_streams(3);
  (void)ctx->get_device_buffer_a().stream_buffer(0);
  (void)ctx->get_gpu_context().get_stream(0);
  (void)ctx2->get_max_tile_sizes();
  gpu::device_vector<T> v(10);
  gpu::copy_to_host(v.data(), a, 10);
}
int main() {
  use<float>(); use<double>(); use<std::complex<float>>(); use<std::complex<double>>();
  return 0;
}
"""


def test_cxx_api_compiles_and_links(tmp_path):
    src = tmp_path / "api.cpp"
    src.write_text(CXX_PROGRAM)
    subprocess.run(["g++", "-std=c++14", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", str(src), "-o", str(tmp_path / "api"), "-L", tiledmm.LIB_DIR,
                    "-ltiledmm", "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{tiledmm.LIB_DIR}"], check=True)
    assert (tmp_path / "api").exists()
