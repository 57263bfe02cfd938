// Workgroup- and grid-wide byte helpers of the charm:: interface
// (reference include/kernel.h:6-28).  The vector copy used on every send is
// charm::rt::wg_copy (16-byte words with an alignment-tiered fallback),
// written for this runtime.
#ifndef _KERNEL_H_
#define _KERNEL_H_

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace charm {

__device__ __forceinline__ void memset_kernel_block(void* addr, int val,
                                                    size_t size) {
  for (size_t i = threadIdx.x; i < size; i += blockDim.x)
    ((char*)addr)[i] = (char)val;
}

__device__ __forceinline__ void memset_kernel_grid(void* addr, int val,
                                                   size_t size) {
  for (size_t i = blockDim.x * blockIdx.x + threadIdx.x; i < size;
       i += gridDim.x * blockDim.x)
    ((char*)addr)[i] = (char)val;
}

__device__ __forceinline__ void memcpy_kernel_block(void* dst, void* src,
                                                    size_t size) {
  for (size_t i = threadIdx.x; i < size; i += blockDim.x)
    ((char*)dst)[i] = ((char*)src)[i];
}

__device__ __forceinline__ void memcpy_kernel_grid(void* dst, void* src,
                                                   size_t size) {
  for (size_t i = blockDim.x * blockIdx.x + threadIdx.x; i < size;
       i += gridDim.x * blockDim.x)
    ((char*)dst)[i] = ((char*)src)[i];
}

}  // namespace charm

#endif  // _KERNEL_H_
