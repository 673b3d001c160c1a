/*
 * gpulda_abi.h -- the drop-in C ABI of libGPUPolyaUrnLDA.so (MI355X / gfx950 build).
 *
 * These five symbols and the two structs are exactly what the reference's Rust host binds:
 *   Args    <- src/ffi.rs:4-13       (== src/cu/train.cuh:13-21)
 *   Buffer  <- src/buffer.rs:4-19    (== src/cu/train.cuh:23-37)
 *   initialize / sample_phi / sample_z_async / cleanup / sync_buffer
 *           <- src/ffi.rs:15-31 (imports by link_name), declared in src/cu/train.cuh:41-45,
 *              implemented in the reference at src/cu/train.cu:31,137,158,98,174.
 *
 * Byte layout (x86-64 SysV, both sides repr(C)):
 *   Args   (32 B): 0 f32 alpha | 4 f32 beta | 8 u32 K | 12 u32 V | 16 u32* C | 24 u32 buffer_size
 *                  | 28 u32 max_D
 *   Buffer (96 B): 0 z* | 8 w* | 16 d* | 24 K_d* | 32 u32 n_docs | 36 u32 n_tokens | 40 gpu_z
 *                  | 48 gpu_w | 56 gpu_d_len | 64 gpu_d_idx | 72 gpu_K_d | 80 gpu_rng | 88 stream
 *
 * Contract (same as the reference):
 *   - No return codes.  A runtime (HIP) error or an invalid argument prints a message to stdout
 *     and terminates the process with exit(code), as the reference's error.cuh:13-18 does.
 *   - Single host thread, not re-entrant.  The library keeps the Args* passed to initialize()
 *     and reads it until cleanup(); the caller keeps Args and C alive until then.
 *   - initialize() allocates device state into the caller's Buffer structs (the gpu_* fields and
 *     stream are library-owned opaque handles; gpu_z stays NULL because z is staged through the
 *     library's pinned host memory as u16); cleanup() frees them.
 *   - sample_phi() is synchronous.  sample_z_async() returns before the sweep finishes; the caller
 *     must not touch buffer->z until sync_buffer(buffer) returns.
 *   - Host arrays w, d and K_d are read-only to the library (train.rs maps them read-only); z is
 *     read and then overwritten with the new topic indicators.
 *
 * Semantic differences from the reference (each a fix of a reference defect) are listed in
 * gpulda_ext.h next to the extension calls that expose them.
 */
#ifndef GPULDA_ABI_H
#define GPULDA_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Args {
  float alpha;          /* Dirichlet prior on theta_d                      (ffi.rs:6)  */
  float beta;           /* Dirichlet prior on phi_k (Poisson offset)       (ffi.rs:7)  */
  uint32_t K;           /* number of topics                                (ffi.rs:8)  */
  uint32_t V;           /* vocabulary size = lines of c.txt                (ffi.rs:9)  */
  uint32_t* C;          /* per-word corpus counts, length V                (ffi.rs:10) */
  uint32_t buffer_size; /* max tokens per buffer                           (ffi.rs:11) */
  uint32_t max_D;       /* max documents per buffer                        (ffi.rs:12) */
} Args;

typedef struct Buffer {
  uint32_t* z;          /* topic indicators, n_tokens (read, then written back) */
  uint32_t* w;          /* word ids, n_tokens (read-only)                        */
  uint32_t* d;          /* document lengths, n_docs (read-only)                  */
  uint32_t* K_d;        /* distinct-topic hints, n_docs (read-only, unused)      */
  uint32_t n_docs;
  uint32_t n_tokens;
  void* gpu_z;          /* library-owned handles from here on */
  void* gpu_w;
  void* gpu_d_len;
  void* gpu_d_idx;
  void* gpu_K_d;
  void* gpu_rng;
  void* stream;
} Buffer;

#ifdef __cplusplus
static_assert(sizeof(Args) == 32, "Args must be 32 bytes (ffi.rs:4-13)");
static_assert(__builtin_offsetof(Args, C) == 16, "Args.C at 16");
static_assert(__builtin_offsetof(Args, buffer_size) == 24, "Args.buffer_size at 24");
static_assert(__builtin_offsetof(Args, max_D) == 28, "Args.max_D at 28");
static_assert(sizeof(Buffer) == 96, "Buffer must be 96 bytes (buffer.rs:4-19)");
static_assert(__builtin_offsetof(Buffer, n_docs) == 32, "Buffer.n_docs at 32");
static_assert(__builtin_offsetof(Buffer, n_tokens) == 36, "Buffer.n_tokens at 36");
static_assert(__builtin_offsetof(Buffer, gpu_z) == 40, "Buffer.gpu_z at 40");
static_assert(__builtin_offsetof(Buffer, stream) == 88, "Buffer.stream at 88");
#endif

/* Replaces src/cu/train.cu:31 (declared train.cuh:41, bound ffi.rs:17-18). */
void initialize(Args* args, Buffer* buffers, uint32_t n_buffers);
/* Replaces src/cu/train.cu:137 (declared train.cuh:42, bound ffi.rs:20-21). */
void sample_phi(void);
/* Replaces src/cu/train.cu:158 (declared train.cuh:43, bound ffi.rs:23-24). */
void sample_z_async(Buffer* buffer);
/* Replaces src/cu/train.cu:98 (declared train.cuh:44, bound ffi.rs:26-27). */
void cleanup(Buffer* buffers, uint32_t n_buffers);
/* Replaces src/cu/train.cu:174 (declared train.cuh:45, bound ffi.rs:29-30). */
void sync_buffer(Buffer* buffer);

#ifdef __cplusplus
}
#endif

#endif /* GPULDA_ABI_H */
