// Logging macros and small helpers of the charm:: interface
// (reference include/common.h:7-13, 87-88).
#ifndef _COMMON_H_
#define _COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <new>  // placement new

#define PINFO(...) printf("[INFO] " __VA_ARGS__)
#define PERROR(...) printf("[ERROR] " __VA_ARGS__)
#ifdef DEBUG
#define PDEBUG(...) printf("[DEBUG] " __VA_ARGS__)
#else
#define PDEBUG(...) \
  do {              \
  } while (0)
#endif

#ifndef SM_LEVEL
// This runtime always maps one PE to one workgroup (the reference's
// SM_LEVEL mode, charming.cu:85-92); examples select that branch.
#define SM_LEVEL 1
#endif

#endif  // _COMMON_H_
