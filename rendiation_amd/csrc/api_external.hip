// api_external.hip -- buffers shared with another graphics API (include/
// rdrt.h: rdrt_external_buffer_*).  The reference keeps its rays, hits and
// images in wgpu buffers (shader/ray-tracing/src/api/backend.rs:48-56 hands
// the encoder's buffers to the ray-tracing pipeline); a backend built on
// this library imports such a buffer, exported by the Vulkan device as an
// opaque file descriptor, and passes its device address to the rdrt_* calls
// instead of copying through host memory.

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cstring>

#include "context.h"

using namespace rdrt;

struct rdrt_external_buffer {
  rdrt_ctx *ctx = nullptr;
  hipExternalMemory_t mem = nullptr;
  void *ptr = nullptr;
};

extern "C" {

int32_t rdrt_external_buffer_import(rdrt_ctx *c, int fd, uint64_t size, uint64_t offset, uint64_t length,
                                    rdrt_external_buffer **out, void **dev_ptr) {
  if (!c || !out || !dev_ptr) return fail(RDRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  *dev_ptr = nullptr;
  if (fd < 0) return fail(RDRT_ERR_INVALID_ARG, "invalid file descriptor");
  if (size == 0 || length == 0 || offset > size || length > size - offset)
    return fail(RDRT_ERR_INVALID_ARG, "the mapped range must lie inside the exported allocation");
  int32_t rc = need_device(c);
  if (rc) return rc;
  hipExternalMemoryHandleDesc hd;
  std::memset(&hd, 0, sizeof(hd));
  hd.type = hipExternalMemoryHandleTypeOpaqueFd;
  hd.handle.fd = fd;
  hd.size = size;
  hipExternalMemory_t mem = nullptr;
  RDRT_HIP(hipImportExternalMemory(&mem, &hd));  // on success the runtime owns the descriptor
  hipExternalMemoryBufferDesc bd;
  std::memset(&bd, 0, sizeof(bd));
  bd.offset = offset;
  bd.size = length;
  void *ptr = nullptr;
  hipError_t e = hipExternalMemoryGetMappedBuffer(&ptr, mem, &bd);
  if (e != hipSuccess) {
    (void)hipDestroyExternalMemory(mem);
    return fail(RDRT_ERR_HIP, std::string("hipExternalMemoryGetMappedBuffer: ") + hipGetErrorString(e));
  }
  rdrt_external_buffer *b = new rdrt_external_buffer();
  b->ctx = c;
  b->mem = mem;
  b->ptr = ptr;
  *out = b;
  *dev_ptr = ptr;
  return RDRT_OK;
}

int32_t rdrt_external_buffer_release(rdrt_external_buffer *b) {
  if (!b) return RDRT_OK;
  (void)hipSetDevice(b->ctx->cfg.device);
  (void)hipStreamSynchronize(b->ctx->work_stream);  // the context's work on it has finished
  (void)hipFree(b->ptr);
  hipError_t e = hipDestroyExternalMemory(b->mem);
  delete b;
  if (e != hipSuccess) return fail(RDRT_ERR_HIP, std::string("hipDestroyExternalMemory: ") + hipGetErrorString(e));
  return RDRT_OK;
}

}  // extern "C"
