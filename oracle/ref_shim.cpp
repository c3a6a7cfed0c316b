// extern "C" entry points into the reference Tiled-MM library, compiled from the reference's
// own sources (see Makefile, target `ref`). TEST INFRASTRUCTURE ONLY: used by the tests to
// pin the oracle's tile policy and geometry and, on a GPU, to run the reference's gemm on the
// same inputs as the library under test.
#include <Tiled-MM/tiled_matrix.hpp>
#include <Tiled-MM/tiled_mm.hpp>
#include <Tiled-MM/util.hpp>

#include <complex>
#include <exception>
#include <iostream>

namespace gpu {
// defined in src/Tiled-MM/mm_handle.cpp, not declared in any header
int optimal_tile_size(int dim, int curr_tile_size, int max_tile_size);
}  // namespace gpu

extern "C" {

int ref_optimal_tile_size(int dim, int curr, int max) { return gpu::optimal_tile_size(dim, curr, max); }

// Geometry of a rows x cols matrix (leading dimension ld) tiled with (tr x tc) tiles:
// out = {num_tiles_row, num_tiles_col, tile_rows(ti,tj), tile_cols(ti,tj), tile_offset(ti,tj)}
void ref_tile_geometry(int rows, int cols, int ld, int tr, int tc, int ti, int tj, long long* out) {
  gpu::tiled_matrix<double> t(nullptr, rows, cols, ld, gpu::tile_dim(tr, tc));
  gpu::tile_dim d = t.tile_dimensions(gpu::tile_coord(ti, tj));
  out[0] = t.num_tiles_row();
  out[1] = t.num_tiles_col();
  out[2] = d.rows();
  out[3] = d.cols();
  out[4] = t.tile_offset(gpu::tile_coord(ti, tj));
}

// The reference's host-to-host gemm (needs a GPU). pin_host_buffers is always false, as in the
// reference's own test (tests/test-multiply.cpp:311,334). With copy_c_back == 0 the resident
// device C (ld = m) is copied into C afterwards.
int ref_dgemm(char ta, char tb, int m, int n, int k, double alpha, double* a, int lda, double* b, int ldb, double beta,
              double* c, int ldc, int streams, int tm, int tn, int tk, int copy_back) {
  try {
    auto ctx = gpu::make_context<double>(streams, tm, tn, tk);
    gpu::gemm(*ctx, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, false, copy_back != 0);
    if (!copy_back) gpu::copy_to_host(ctx->get_full_device_buffer_c().data(), c, (size_t)m * n);
    return 0;
  } catch (std::exception& e) {
    std::cerr << "reference gemm failed: " << e.what() << std::endl;
    return 1;
  }
}

int ref_zgemm(char ta, char tb, int m, int n, int k, const double* alpha, std::complex<double>* a, int lda,
              std::complex<double>* b, int ldb, const double* beta, std::complex<double>* c, int ldc, int streams,
              int tm, int tn, int tk, int copy_back) {
  try {
    auto ctx = gpu::make_context<std::complex<double>>(streams, tm, tn, tk);
    gpu::gemm(*ctx, ta, tb, m, n, k, std::complex<double>(alpha[0], alpha[1]), a, lda, b, ldb,
              std::complex<double>(beta[0], beta[1]), c, ldc, false, copy_back != 0);
    if (!copy_back) gpu::copy_to_host(ctx->get_full_device_buffer_c().data(), c, (size_t)m * n);
    return 0;
  } catch (std::exception& e) {
    std::cerr << "reference gemm failed: " << e.what() << std::endl;
    return 1;
  }
}

}  // extern "C"
