// C++ drop-in check of the stream, event and buffer surface a caller of the reference can reach
// through mm_handle (Tiled-MM src/Tiled-MM/gpu_context.cpp:6-57, device_stream.hpp:9-86,
// device_event.hpp:9-77, device_buffer.hpp:40-83): stream access and its range checks, events
// enqueued on one stream and waited on by another, event timing, the per-stream tile slots, and
// a stream count raised after creation, followed by a GEMM on the grown context. Exit status 0
// and "CONTEXT OK" on success; every failed check prints one line.
#include <Tiled-MM/tiled_mm.hpp>
#include <Tiled-MM/util.hpp>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

template <typename F>
bool throws(F&& f) {
  try {
    f();
  } catch (const std::runtime_error&) {
    return true;
  }
  return false;
}

}  // namespace

int main() {
  auto ctx = gpu::make_context<double>(2, 512, 512, 512);
  gpu::gpu_context& gc = ctx->get_gpu_context();

  // streams and their range checks (gpu_context.cpp:21-37)
  expect(gc.get_num_streams() == 2, "two streams after make_context(2, ...)");
  hipStream_t s0 = gc.get_stream(0), s1 = gc.get_stream(1);
  expect(s0 != nullptr && s1 != nullptr && s0 != s1, "distinct compute streams 0 and 1");
  expect(gc.get_device_stream(0).stream() == s0 && gc.get_device_stream(1).stream() == s1,
         "get_device_stream(i) wraps get_stream(i)");
  hipStream_t r = gc.get_result_stream().stream();
  expect(r != nullptr && r != s0 && r != s1, "result stream distinct from the compute streams");
  expect(throws([&] { gc.get_stream(2); }), "get_stream(n_streams) throws");
  expect(throws([&] { gc.get_stream(-1); }), "get_stream(-1) throws");
  expect(throws([&] { gc.get_device_stream(5); }), "get_device_stream(5) throws");
  expect(throws([&] { gc.enqueue_event(2); }), "enqueue_event(n_streams) throws");

  // events: a copy on stream 0, an event after it, stream 1 waits for the event and copies the
  // result on; the data stream 1 reads must be the data stream 0 wrote
  const size_t n = (size_t)16 << 20;  // 128 MiB of doubles
  double* host = gpu::malloc_pinned<double>(n, 0.0);
  for (size_t i = 0; i < n; ++i) host[i] = (double)(i % 1000003);
  gpu::device_vector<double> x(n), y(n);
  gpu::device_event start = gc.enqueue_event(0);
  gpu::copy_to_device_async(host, x.data(), n, s0);
  gpu::device_event written = gc.get_device_stream(0).enqueue_event();
  gc.get_device_stream(1).wait_on_event(written);
  gpu::check_runtime_status(hipMemcpyAsync(y.data(), x.data(), n * sizeof(double), hipMemcpyDeviceToDevice, s1));
  gpu::device_event copied = gc.enqueue_event(1);
  copied.wait();
  written.wait();
  const double secs = written.time_since(start);
  expect(secs > 0.0 && secs < 10.0, "time_since between two events of stream 0 is positive");
  std::vector<double> back(n);
  gpu::copy_to_host(y.data(), back.data(), n);
  bool same = true;
  for (size_t i = 0; i < n && same; ++i) same = back[i] == host[i];
  expect(same, "stream 1 saw stream 0's copy after waiting on its event");

  // per-stream tile slots (device_buffer.hpp:44-54): A tiles are tile_m x tile_k
  ctx->set_tile_sizes(256, 128, 64);
  gpu::device_buffer<double>& ab = ctx->get_device_buffer_a();
  double* a0 = ab.stream_buffer(0);
  double* a1 = ab.stream_buffer(1);
  expect(a0 != nullptr && a1 == a0 + 256 * 64, "stream slots of buffer A are tile_m * tile_k apart");
  expect(throws([&] { ab.stream_buffer(2); }), "stream_buffer(n_streams) throws");
  expect(throws([&] { ctx->get_device_buffer_c().stream_buffer(-1); }), "stream_buffer(-1) throws");
  gpu::copy_to_device(host, a1, 256 * 64);
  std::vector<double> slot(256 * 64);
  gpu::copy_to_host(a1, slot.data(), slot.size());
  expect(std::memcmp(slot.data(), host, slot.size() * sizeof(double)) == 0, "a tile slot holds what was copied in");
  double* c1 = ctx->get_device_buffer_c().stream_buffer(1);
  double* c0 = ctx->get_device_buffer_c().stream_buffer(0);
  expect(c1 == c0 + 256 * 128, "stream slots of buffer C are tile_m * tile_n apart");

  // more streams after creation (gpu_context.cpp:49-53)
  ctx->set_num_streams(3);
  expect(gc.get_num_streams() == 3 && ctx->get_num_streams() == 3, "three streams after set_num_streams(3)");
  hipStream_t s2 = gc.get_stream(2);
  expect(s2 != nullptr && s2 != gc.get_stream(0) && s2 != gc.get_stream(1), "a third, distinct stream");
  expect(throws([&] { gc.get_stream(3); }), "get_stream(3) throws with three streams");
  expect(ctx->get_device_buffer_b().stream_buffer(2) != nullptr, "buffer B has a slot for stream 2");

  // a GEMM on the grown context, integer data: exact
  const int m = 700, nn = 600, k = 500;
  std::vector<double> A((size_t)m * k), B((size_t)k * nn), C((size_t)m * nn), ref((size_t)m * nn);
  for (size_t i = 0; i < A.size(); ++i) A[i] = (double)((i * 7) % 10);
  for (size_t i = 0; i < B.size(); ++i) B[i] = (double)((i * 3) % 10);
  for (size_t i = 0; i < C.size(); ++i) C[i] = (double)(i % 10);
  for (int j = 0; j < nn; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0;
      for (int p = 0; p < k; ++p) s += A[i + (size_t)p * m] * B[p + (size_t)j * k];
      ref[i + (size_t)j * m] = 2.0 * s + C[i + (size_t)j * m];
    }
  ctx->set_tile_sizes(256, 256, 256);
  gpu::gemm(*ctx, 'N', 'N', m, nn, k, 2.0, A.data(), m, B.data(), k, 1.0, C.data(), m, false, true);
  expect(C == ref, "GEMM on the context with three streams is exact");

  gpu::free_pinned(host);
  std::printf(failures ? "CONTEXT FAILED (%d)\n" : "CONTEXT OK\n", failures);
  return failures ? 1 : 0;
}
