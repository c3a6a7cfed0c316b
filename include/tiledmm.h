/*
 * tiledmm.h - C ABI of the MI355X-native out-of-core GEMM library (libtiledmm.so).
 *
 * Every entry point returns a tmm_status (0 = success). On failure the handle's message is
 * available through tmm_last_error(h) (or tmm_last_error(NULL) for errors without a handle).
 * Matrices are column-major. Complex scalars and matrices are interleaved (re, im) pairs,
 * binary compatible with std::complex<T> (the reference relies on the same layout,
 * Tiled-MM src/Tiled-MM/tiled_mm.cpp:219-223).
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * Tiled-MM source tree).
 */
#ifndef TILEDMM_H
#define TILEDMM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tmm_handle tmm_handle;

typedef enum {
  TMM_F32 = 0, /* float                */
  TMM_F64 = 1, /* double               */
  TMM_C32 = 2, /* std::complex<float>  */
  TMM_C64 = 3  /* std::complex<double> */
} tmm_dtype;

typedef enum {
  TMM_SUCCESS = 0,
  TMM_ERR_INVALID_ARG = 1, /* bad dimension, leading dimension, op flag, stream id, tile size */
  TMM_ERR_HIP = 2,         /* a HIP runtime call or kernel launch failed                      */
  TMM_ERR_ALLOC = 3,       /* device or pinned host allocation failed                         */
  TMM_ERR_DTYPE = 4,       /* entry point called on a handle of another element type           */
  TMM_ERR_INTERNAL = 5
} tmm_status;

/* Handle creation: replaces gpu::make_context<T>(streams, tile_m, tile_n, tile_k)
 * (mm_handle.hpp:60-76) and the mm_handle<T> constructor (mm_handle.cpp:10-29). Streams and
 * buffers belong to the device current at creation time. */
int tmm_create(tmm_handle** out, int dtype, int n_streams, int max_tile_m, int max_tile_n, int max_tile_k);
/* mm_handle<T>::~mm_handle (mm_handle.cpp:31-34): frees streams and device buffers. */
int tmm_destroy(tmm_handle* h);
int tmm_get_dtype(tmm_handle* h, int* dtype);

/* mm_handle<T>::set_num_streams / get_num_streams (mm_handle.cpp:36-50). */
int tmm_set_num_streams(tmm_handle* h, int n_streams);
int tmm_get_num_streams(tmm_handle* h, int* n_streams);
/* mm_handle<T>::set_tile_sizes (mm_handle.cpp:57-71): tiles must not exceed the maxima. */
int tmm_set_tile_sizes(tmm_handle* h, int tile_m, int tile_n, int tile_k);
/* Current tile sizes (device_buffer<T>::get_tile_sizes, device_buffer.hpp:73-76). */
int tmm_get_tile_sizes(tmm_handle* h, int* tile_m, int* tile_n, int* tile_k);
/* mm_handle<T>::get_max_tile_sizes (mm_handle.cpp:112-115). */
int tmm_get_max_tile_sizes(tmm_handle* h, int* tile_m, int* tile_n, int* tile_k);
/* mm_handle<T>::set_streams_and_tiles (mm_handle.cpp:135-145). */
int tmm_set_streams_and_tiles(tmm_handle* h, int n_streams, int tile_m, int tile_n, int tile_k);
/* mm_handle<T>::set_full_sizes (mm_handle.cpp:72-80): size the resident C to m*n. */
int tmm_set_full_sizes(tmm_handle* h, int m, int n, int k);
/* mm_handle<T>::optimal_tile_sizes (mm_handle.cpp:117-133) with optimal_tile_size (:89-110). */
int tmm_optimal_tile_sizes(tmm_handle* h, int m, int n, int k, int* tile_m, int* tile_n, int* tile_k);
/* The reference policy for one dimension, without a handle (mm_handle.cpp:89-110). */
int tmm_optimal_tile_size(int dim, int curr_tile_size, int max_tile_size);

/* Host-to-host GEMM: C = alpha*op(A)*op(B) + beta*C with host pointers. Replaces
 * gpu::gemm<T> (tiled_mm.hpp:69-79, tiled_mm.cpp:492-624). op in 'N','T','C' (lower case
 * accepted, any other letter means 'N' as in tiled_mm.cpp:168-179). alpha and beta point at
 * one scalar (two reals for complex). pin_host_buffers: accepted for API compatibility;
 * pageable host memory is always streamed through the library's pinned staging buffers
 * instead of being page-locked per call. copy_c_back = 0 leaves the result in the resident
 * device C (ld = m), see tmm_full_c_device_ptr. Synchronous: returns when C is final. */
int tmm_sgemm(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const float* alpha,
              const float* a, int ld_a, const float* b, int ld_b, const float* beta, float* c, int ld_c,
              int pin_host_buffers, int copy_c_back);
int tmm_dgemm(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const double* alpha,
              const double* a, int ld_a, const double* b, int ld_b, const double* beta, double* c, int ld_c,
              int pin_host_buffers, int copy_c_back);
int tmm_cgemm(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const float* alpha,
              const float* a, int ld_a, const float* b, int ld_b, const float* beta, float* c, int ld_c,
              int pin_host_buffers, int copy_c_back);
int tmm_zgemm(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const double* alpha,
              const double* a, int ld_a, const double* b, int ld_b, const double* beta, double* c, int ld_c,
              int pin_host_buffers, int copy_c_back);

/* Batched host-to-host FP64 GEMM: count independent problems C_i = alpha_i*op(A_i)*op(B_i) +
 * beta_i*C_i of uneven sizes, one op pair for all, host pointers per problem. No counterpart in
 * the reference (there: one gpu::gemm call per problem, tiled_mm.cpp:492-624). The batch is packed
 * into one device arena, cut into groups whose H2D copies, one persistent-kernel launch and D2H
 * copies overlap on the copy and compute queues, and (with pinned host memory) may be captured
 * into one hipGraph. The library's autotuner (rates measured on the device, optionally cached in
 * the file named by TMM_TUNE_FILE) chooses the group count, the persistent grid, the tile body
 * (register-staged or direct-to-LDS), the compute queue count, graph or queues, and, when the
 * handle has TMM_OPT_OZAKI_MODULI set, which problems run as Ozaki scheme II products (those
 * whose modelled emulated time is shorter). Requires a TMM_F64 handle. Synchronous. */
int tmm_dgemm_batched(tmm_handle* h, char trans_a, char trans_b, int count, const int* m, const int* n, const int* k,
                      const double* alpha, const double* const* a, const int* ld_a, const double* const* b,
                      const int* ld_b, const double* beta, double* const* c, const int* ld_c);
/* Configuration the autotuner chose for the last batched call, and the measured rates
 * (rates[0] H2D B/s, [1] D2H B/s, [2] persistent-kernel flop/s, [3] per-group overhead s). */
int tmm_batch_config(tmm_handle* h, int* groups, int* workgroups, int* graph, double* predicted_s, double* rates);
/* All rates the autotuner has measured on the handle (n <= 7): [0] H2D B/s, [1] D2H B/s, [2]
 * register-staged persistent kernel flop/s, [3] per-group overhead s, [4] direct-to-LDS persistent
 * kernel flop/s, [5] Ozaki scheme II FP64-equivalent flop/s at the handle's moduli (0 if not
 * measured), [6] fixed seconds of one emulated product. */
int tmm_batch_rates(tmm_handle* h, double* rates, int n);

/* mm_handle<T>::get_full_device_buffer_c (mm_handle.cpp:162-165): device pointer and element
 * count of the resident C (column-major, ld = m) after a copy_c_back = 0 call. */
int tmm_full_c_device_ptr(tmm_handle* h, void** ptr, size_t* n_elems);
/* mm_handle<T>::get_device_buffer_{a,b,c} + device_buffer<T>::stream_buffer
 * (mm_handle.cpp:147-160, device_buffer.hpp:44-54): per-stream tile slot of operand
 * which = 0 (A), 1 (B), 2 (C), allocated on first request with the current tile sizes. */
int tmm_tile_buffer(tmm_handle* h, int which, int stream_id, void** ptr, size_t* n_elems);
/* gpu_context::get_stream / get_result_stream (gpu_context.cpp:27-32,55-57). The handle
 * owns n_streams compute streams and one result (device-to-host) stream; stream_id = -1
 * returns the result stream. */
int tmm_get_stream(tmm_handle* h, int stream_id, void** stream);

/* Message of the last failure on h (h = NULL: last failure without a handle). The
 * reference prints to std::cerr and throws std::runtime_error("GPU ERROR"),
 * src/Tiled-MM/util.hpp:13-36. */
const char* tmm_last_error(tmm_handle* h);

/* Device-resident tile GEMM, the per-tile kernel on its own: replaces the vendor BLAS call
 * cublas_gemm_wrapper -> {s,d,c,z}gemm (tiled_mm.cpp:181-268). All pointers are device
 * pointers; stream = NULL runs on the handle's first compute stream. Asynchronous. */
int tmm_device_gemm(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const void* alpha,
                    const void* a, int ld_a, const void* b, int ld_b, const void* beta, void* c, int ld_c,
                    void* stream);
/* tmm_device_gemm without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, semantics and kernels as tmm_device_gemm; with an Ozaki option the
 * emulation works in a fixed scratch array of TMM_NOALLOC_SCRATCH_BYTES (256 MiB) that the
 * library holds on every device it runs on until the process ends (its address is looked up by
 * tmm_create) instead of the handle's workspace. A product whose workspace would exceed it is
 * refused with TMM_ERR_INVALID_ARG (16 moduli: up to about 2048^3). All calls on one device share
 * the scratch: calls that overlap on one device (other streams, other handles) race, so callers
 * order them, e.g. on one stream. Asynchronous. */
#define TMM_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int tmm_device_gemm_noalloc(tmm_handle* h, char trans_a, char trans_b, int m, int n, int k, const void* alpha,
                            const void* a, int ld_a, const void* b, int ld_b, const void* beta, void* c, int ld_c,
                            void* stream);
/* Wait for all work the handle has queued (the reference's gemm ends with a device-wide
 * DeviceSynchronize, tiled_mm.cpp:602-604; this waits for the handle's streams only). */
int tmm_synchronize(tmm_handle* h);

/* Tuning switches with no counterpart in the reference.
 *   TMM_OPT_COARSEN      1 (default): small reference tiles are merged into larger blocks so
 *                        one tile GEMM fills the GPU, block edges follow the tile kernel's
 *                        workgroup grid (same block count, each a multiple of the workgroup
 *                        tile but the last), and Ozaki scheme II calls merge k-chunks;
 *                        0 keeps the reference tile grid.
 *   TMM_OPT_FIRST_CHUNK  k-size of the first k-chunk (0 = same as the others, -1 = auto):
 *                        a short first chunk lets the first tile GEMM start sooner.
 *   TMM_OPT_HOST_THREADS threads used to stage pageable host memory (0 = auto).
 *   TMM_OPT_POOL_LIMIT_MB device work-pool budget in MiB (0 = auto: 85% of free memory).
 *   TMM_OPT_PROFILE      1: time every tile GEMM launch with HIP events recorded on the stream
 *                        it runs on (see tmm_last_call_stats).
 *   TMM_OPT_OZAKI_SLICES 0 (default): FP64 tile GEMMs on FP64 matrix cores. s in [1, 12]: FP64
 *                        GEMMs (tmm_dgemm, tmm_device_gemm on a TMM_F64 handle) run as an
 *                        Ozaki split-precision product: s int8 slices of 7 bits per element
 *                        after per-row / per-column power-of-two scaling, s(s+1)/2 INT8 MFMA
 *                        products with exact int32 accumulation, FP64 recombination.
 *   TMM_OPT_OZAKI_MODULI 0 (default) or N in [6, 20]: FP64 GEMMs run as Ozaki scheme II:
 *                        entries scaled per row / column by powers of two and rounded to
 *                        integers, their product recovered from N INT8 MFMA products modulo
 *                        pairwise coprime moduli <= 256 (CRT reconstruction in FP64).
 *                        Takes precedence over TMM_OPT_OZAKI_SLICES. Inputs must be finite.
 *   TMM_OPT_OZAKI_REUSE  1 (default): with TMM_OPT_OZAKI_MODULI in host-to-host calls, the
 *                        exponents and residues of an A or B block read by several tile
 *                        products are computed once per call and kept in device memory (at
 *                        most half of the free memory, 64 GB); 0: once per tile product. The
 *                        results are the same bit for bit.
 *   TMM_OPT_BATCH_GRAPH  tmm_dgemm_batched with pinned host matrices: -1 (default) the autotuner
 *                        chooses between capturing the call's copies and launches into one
 *                        hipGraph (no per-group host overhead; the graph runs its copies one
 *                        after the other) and enqueueing them on the three queues (H2D and D2H
 *                        overlap); 1 always the graph; 0 never.
 *   TMM_OPT_BATCH_BODY   tmm_dgemm_batched tile body: -1 (default) the autotuner's choice (the faster
 *                        one measured), 0 register-staged LDS double buffer, 1 direct-to-LDS ring.
 *   TMM_OPT_BATCH_QUEUES tmm_dgemm_batched compute queues: 0 (default) the autotuner's choice, 1 or 2.
 *   TMM_OPT_BATCH_EMULATION tmm_dgemm_batched on a handle with TMM_OPT_OZAKI_MODULI set: -1 (default)
 *                        each problem runs as Ozaki scheme II products when the autotuner's model
 *                        says that is faster than native FP64, 0 never, 1 always.
 *   TMM_OPT_TILE_GROUPS  host-to-host calls (all element types, with TMM_OPT_DGEMM_KERNEL 1; FP32
 *                        groups take tiles of more than 128 rows): 1 = the tile GEMMs of each
 *                        steady k-chunk (one whose preceding chunk is at least as long, not the
 *                        last) run as grouped launches of up to 16 tiles (one partial last wave
 *                        per group instead of per tile); v >= 2 (default 2) = as 1, and the other
 *                        chunks after the first also run in groups of at most v tiles, the
 *                        queue's last tile alone; 0 = one launch per tile. Same results.
 *   TMM_OPT_DEFER_CIN    Ozaki scheme II host-to-host calls on pinned host memory: the copy of a
 *                        C tile (read only by the reconstruction of its last k-chunk) is issued
 *                        after the A / B copies of the tile this many places later (default 1),
 *                        and the tile's products start before it arrives; 0: before the tile's
 *                        own products, as for native calls. Same results.
 *   TMM_OPT_DGEMM_KERNEL tile kernels of every element type: 1 (default) = direct global->LDS
 *                        loads into a 3-stage ring (with at least two workgroup tiles per CU
 *                        software-pipelined; FP64 on 128 x 128 tiles, or 128 x 256 from 2^27
 *                        elements of C on); 0 = LDS double buffer filled through registers
 *                        (the kernel that variant 1 also uses for shapes it does not
 *                        take: K % 8 remainders (FP32: K % 16); an odd row count of a
 *                        row-contiguous operand (FP32: not a multiple of 4); for FP64 also m or
 *                        n < 2). */
enum {
  TMM_OPT_COARSEN = 1,
  TMM_OPT_FIRST_CHUNK = 2,
  TMM_OPT_HOST_THREADS = 3,
  TMM_OPT_POOL_LIMIT_MB = 4,
  TMM_OPT_PROFILE = 5,
  TMM_OPT_OZAKI_SLICES = 6,
  TMM_OPT_DGEMM_KERNEL = 7,
  TMM_OPT_OZAKI_MODULI = 8,
  TMM_OPT_OZAKI_REUSE = 9,
  TMM_OPT_BATCH_GRAPH = 10,
  TMM_OPT_TILE_GROUPS = 11,
  TMM_OPT_DEFER_CIN = 12,
  TMM_OPT_BATCH_BODY = 13,
  TMM_OPT_BATCH_QUEUES = 14,
  TMM_OPT_BATCH_EMULATION = 15
};
int tmm_set_option(tmm_handle* h, int option, long value);
int tmm_get_option(tmm_handle* h, int option, long* value);

/* Statistics of the last host-to-host call, for measurement:
 *   stats[0] h2d bytes, [1] d2h bytes, [2] tile GEMM launches, [3] tile_m, [4] tile_n,
 *   [5] tile_k, [6] k-chunks, [7] wall seconds of the call, [8] 1 if A,B,C were all pinned;
 *   with TMM_OPT_PROFILE: [9] summed tile-GEMM kernel seconds, [10] launches timed, [11] their
 *   flops, [12] longest launch in seconds; [13] INT8 operations of the Ozaki path (2mnk per slice
 *   product); [14] 1 if the last batched call ran as one hipGraph, [15] its group count; with
 *   TMM_OPT_PROFILE, of the timed launches those that ran several tiles in one grouped launch:
 *   [16] their count, [17] their summed seconds, [18] their flops; [19] 1 once the call has started
 *   writing the caller's C on the host (0 after a failure means C is unchanged); the last batched
 *   call's autotuner choices: [20] tile body (0 register-staged, 1 direct-to-LDS), [21] compute
 *   queues, [22] problems run as Ozaki scheme II products; with TMM_OPT_PROFILE: [23] the
 *   algorithmic HBM bytes of the grouped launches (distinct A and B blocks read once, each C tile
 *   read and written once per operand use). n <= 24. */
int tmm_last_call_stats(tmm_handle* h, double* stats, int n);

#ifdef __cplusplus
}
#endif
#endif /* TILEDMM_H */
