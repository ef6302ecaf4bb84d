/*
 * include/cospan_hip.h — libcospan_hip: the thin C boundary between CoSpan's C++ host layer and the HIP runtime.
 *
 * The reference's C++ host layer talks to AdaptiveCpp SYCL and SlabKit directly (queues, events, USM vectors,
 * parallel_for). This build keeps the C++ API (header-only, the naga/cospan headers) and routes every runtime action through
 * the plain-C entry points below: opaque handles, plain pointers and sizes, int status codes, a thread-local message.
 * Nothing here throws; the C++ layer maps a non-zero status to the reference's exception types.
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference checkout).
 */
#ifndef COSPAN_HIP_H
#define COSPAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------------------------------ */
#define CS_HIP_OK 0
#define CS_HIP_ERR_INVALID 1 /* a bad argument (null handle, size out of range, ...) */
#define CS_HIP_ERR_RUNTIME 2 /* the HIP runtime reported an error */
#define CS_HIP_ERR_STATE 3   /* a call out of order (capture not begun, ...) */

typedef struct cs_hip_stream_s* cs_hip_stream;
typedef struct cs_hip_event_s*  cs_hip_event;
typedef struct cs_hip_graph_s*  cs_hip_graph;

/* The last error message of this thread (empty when none). Returns the message length. */
int cs_hip_last_error(char* buf, size_t len);

/* The identity of the sources this library was built from (16 hex digits, cospanmdao_amd/source_hash.py); no reference
 * counterpart (the reference compiles its headers into each user program). Tests compare it with the tree. */
const char* cs_hip_build_id(void);

/* ---- device ------------------------------------------------------------------------------------------------------
 * replaces: naga::cospan::detail::Runtime construction and device selection (cospanmdao/include/naga/cospan/runtime.hpp:33-50)
 */
int cs_hip_init(int device);
int cs_hip_device_count(int* count);
/* name (buf), compute units, LDS bytes per work-group, max threads per work-group */
int cs_hip_device_info(int device, char* name, size_t name_len, int* cus, int* lds_bytes, int* max_wg);
int cs_hip_device_synchronize(void);

/* ---- streams and events ------------------------------------------------------------------------------------------
 * replaces: SlabKit's unordered_queue submission with explicit dependency lists, and the read/write event lists of
 * every container (cospanmdao/include/naga/cospan/variable.hpp:98-172; component.hpp:507-539 gather_deps/record_launch).
 * A launch waits on events (hipStreamWaitEvent) instead of listing sycl::event dependencies.
 */
int cs_hip_stream_create(cs_hip_stream* s);
int cs_hip_stream_destroy(cs_hip_stream s);
int cs_hip_stream_synchronize(cs_hip_stream s);
int cs_hip_event_create(cs_hip_event* e, int timing);
int cs_hip_event_destroy(cs_hip_event e);
int cs_hip_event_record(cs_hip_event e, cs_hip_stream s);
int cs_hip_stream_wait_event(cs_hip_stream s, cs_hip_event e);
/* 1 when complete, 0 when pending */
int cs_hip_event_query(cs_hip_event e, int* done);
int cs_hip_event_synchronize(cs_hip_event e);
int cs_hip_event_elapsed_ms(float* ms, cs_hip_event start, cs_hip_event stop);

/* ---- memory --------------------------------------------------------------------------------------------------------
 * replaces: sk::usm_vector allocation/free through the Runtime's allocator (runtime.hpp:33-50; variable.hpp:446-468) and
 * init::from_host / push_from copies (init.hpp:220-260; variable.hpp:550-600).
 */
int cs_hip_malloc(void** p, size_t bytes);
int cs_hip_free(void* p);
int cs_hip_host_malloc(void** p, size_t bytes); /* pinned host memory */
int cs_hip_host_free(void* p);
int cs_hip_memcpy_async(void* dst, const void* src, size_t bytes, cs_hip_stream s);
/* count 32-bit words set to value with a kernel (never a memset node, so it is safe inside a captured graph) */
int cs_hip_fill_u32_async(void* dst, uint32_t value, size_t count, cs_hip_stream s);
/* bytes set to value with a kernel (same reason) */
int cs_hip_fill_u8_async(void* dst, uint8_t value, size_t count, cs_hip_stream s);

/* ---- kernel launch ------------------------------------------------------------------------------------------------
 * replaces: launch_group / launch_range (cospanmdao/include/naga/cospan/kernel.hpp:1015-1087): the host stub of a
 * hipcc-instantiated fused kernel, its grid (x fastest), block, dynamic LDS and kernel arguments.
 */
int cs_hip_launch(const void* kernel, const uint32_t grid[3], const uint32_t block[3], size_t lds_bytes, void** args,
                  cs_hip_stream s);
/* The most work-groups of `block_threads` threads and `lds_bytes` of LDS that can be resident at once on the device. */
int cs_hip_occupancy(const void* kernel, int block_threads, size_t lds_bytes, int* groups_per_cu, int* cus);

/* ---- graphs ---------------------------------------------------------------------------------------------------------
 * replaces: the reference's per-forward re-description and submission (component.hpp:779-970): a steady-state run is
 * captured once (stream capture, including the event fork/join of side streams) and replayed.
 */
int cs_hip_graph_begin(cs_hip_stream s);
int cs_hip_graph_end(cs_hip_stream s, cs_hip_graph* g);
int cs_hip_graph_launch(cs_hip_graph g, cs_hip_stream s);
int cs_hip_graph_destroy(cs_hip_graph g);
/* 1 while the stream is capturing */
int cs_hip_stream_capturing(cs_hip_stream s, int* capturing);

/* ---- bounded device waits ------------------------------------------------------------------------------------------
 * (no reference counterpart: the reference's ticket waits, cospanmdao/include/naga/cospan/kernel.hpp:197-236, spin
 * without a bound.) Every spin-wait in device code of this build (a fused group's wait_group / global barrier in ticket
 * mode, the scan's look-back, a work-queue kernel's waits) polls at most `limit` times. A wait that gives up ORs a bit
 * below into the fault word of its device and the kernel runs on to its normal end; the results of that work are then
 * not valid. A limit of 0 makes every wait give up without polling (a way for a test to force the condition from the
 * host). The C++ layer's Model::wait and view synchronisations read and clear the word and throw std::runtime_error
 * when it is set; callers of the C entry points read it with cs_hip_take_fault after their work has completed. The word
 * is per device: the first reader after the faulting work takes it.
 */
#define CS_HIP_FAULT_TICKET_WAIT 1u   /* a fused group's wait for an earlier group or segment gave up */
#define CS_HIP_FAULT_SCAN_LOOKBACK 2u /* the scan's look-back at a predecessor tile gave up */
#define CS_HIP_FAULT_STATE 4u         /* a work-group found ticket state or a scan workspace that was not zeroed */
#define CS_HIP_FAULT_UNIT_WAIT 8u     /* a work-queue kernel's wait for an earlier unit gave up */
#define CS_HIP_DEFAULT_SPIN_LIMIT (1u << 24)
/* The process-wide poll bound of device waits launched from now on (CS_HIP_DEFAULT_SPIN_LIMIT at start). */
int      cs_hip_set_spin_limit(uint32_t limit);
uint32_t cs_hip_spin_limit(void);
/* The device address of the current device's fault word (what kernels OR their fault bits into). */
int cs_hip_fault_word(uint32_t** device_word);
/* Reads and clears the current device's fault word; *bits = 0 when no wait gave up. Call after the work completed. */
int cs_hip_take_fault(uint32_t* bits);

/* ---- device exclusive scan ---------------------------------------------------------------------------------------
 * replaces: naga::cospan::detail::exclusive_scan_u32 (cospanmdao/include/naga/cospan/device_scan.hpp:46-110):
 * out[0..n] = the exclusive scan of in[0..n) (n + 1 outputs, out[n] the total), running sums wrap mod 2^32.
 * One launch: a decoupled look-back scan (tickets in start order, wave-parallel look-back with ballot, DPP wave scans,
 * LDS staging). `workspace` must hold cs_hip_scan_workspace_bytes(n) bytes, zeroed once (cs_hip_scan_workspace_init);
 * every scan leaves it zeroed again. Two scans that run at the same time need two workspaces. workspace == NULL uses a
 * library-owned workspace of the current device and of stream `s` (one per stream, so scans on different streams never
 * share one; it grows outside any capture, and an outgrown one is kept until the process ends because a captured graph
 * may still name it). A graph that captured such a scan must be replayed on its capturing stream, or given a workspace.
 * A look-back that gives up sets CS_HIP_FAULT_SCAN_LOOKBACK (cs_hip_take_fault).
 */
size_t cs_hip_scan_workspace_bytes(size_t n);
int    cs_hip_scan_workspace_init(void* workspace, size_t bytes, cs_hip_stream s);
int    cs_hip_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* workspace, cs_hip_stream s);

/* The same scan with no allocation of any kind. No counterpart in the reference: it is for callers that cannot allocate
 * (code under graph capture, tight serving loops, harnesses that forbid hipMalloc). Its workspace is a fixed array the
 * library holds on each device until the process ends (CS_HIP_SCAN_STATIC_BYTES), so n may be at most
 * CS_HIP_SCAN_STATIC_MAX_N (larger n: CS_HIP_ERR_INVALID, nothing launched). The first call on a device clears the array
 * with a kernel on stream s; every scan leaves it zeroed. All calls on one device share the array: calls that overlap on
 * one device (two streams, two threads) race, so keep them on one stream. */
#define CS_HIP_SCAN_STATIC_MAX_N (1ull << 26)
#define CS_HIP_SCAN_STATIC_BYTES (16u + 8u * (unsigned)(CS_HIP_SCAN_STATIC_MAX_N / 2048u))
int cs_hip_exclusive_scan_u32_static(const uint32_t* in, uint32_t* out, size_t n, cs_hip_stream s);

/* ---- ragged column build --------------------------------------------------------------------------------------------
 * replaces: build_column (cospanmdao/include/naga/cospan/shape_hook.hpp:215-348) with its flag bits (:180-201).
 * src_dtype: 0 i32, 1 u32, 2 i64, 3 u64. counts mode (offsets_mode 0): src is [count_src, k] row-major, column 0
 * drives extents[i] = max(v, 0) and offs = their exclusive scan (count_src + 1 entries). offsets mode: src holds
 * count_src = N + 1 offsets, extents[i] = offs[i + 1] - offs[i]. result[0] = the total, result[1] = the flag word
 * (1: negative / invalid, 2: past 32 bits or wrapped, 4: a bounded device wait gave up, so the offsets are not valid);
 * both written by the last work-group to finish (device memory). One launch in either mode.
 */
#define CS_HIP_FLAG_NEGATIVE 1u
#define CS_HIP_FLAG_EXCEEDS32 2u
#define CS_HIP_FLAG_WAIT_TIMEOUT 4u
int cs_hip_build_column(int src_dtype, const void* src, size_t count_src, size_t k, int offsets_mode, uint32_t* extents,
                        uint32_t* offs, uint32_t* result, void* workspace, cs_hip_stream s);

/* ---- initializers on the device ------------------------------------------------------------------------------------
 * replaces: the host-side generators examples/gemver-chain/common/chain_data.hpp:34-63 run as a custom initializer
 * (init::custom, init.hpp:203-329): element (i, j) of a rows x cols row-major array with leading dimension ld gets
 * hash_unit(seed, i * cols + j), bit-identical to the host generator.
 */
int cs_hip_fill_hash_f32(uint32_t seed, float* out, size_t rows, size_t cols, size_t ld, cs_hip_stream s);

/* ---- reduced-precision storage (no reference counterpart: dtype.hpp:17-20 has no bf16 / fp8) ---------------------- */
/* round to nearest even */
int cs_hip_quantize_bf16(const float* in, uint16_t* out, size_t rows, size_t cols, size_t ld_in, size_t ld_out,
                         cs_hip_stream s);
/* per-row scale[i] = max_j |a_ij| / 448 (1 for a zero row), q_ij = e4m3(a_ij / scale[i]), OCP E4M3, nearest even,
 * saturating */
int cs_hip_quantize_e4m3_rows(const float* in, uint8_t* out, float* scale, size_t rows, size_t cols, size_t ld_in,
                              size_t ld_out, cs_hip_stream s);

#ifdef __cplusplus
}
#endif
#endif
