// noalloc.hip - the device scratch of the allocation-free entry points.
//
// One file-scope device array serves every call that must not allocate
// (bn_solver_adjoint_grad_noalloc): each GPU the library runs on holds its
// own copy for the life of the process.  Host code never uses the array's
// name as an address (on the host it is only a stand-in): the first call
// on a device resolves that device's copy with hipGetSymbolAddress, with
// the device current, and every later use goes through the pointer kept
// for that device.  Calls that overlap on one GPU share the one copy and
// race; callers serialise them.
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "bn_amg_ext.h"
#include "bn_kernels.h"

static __device__ __align__(256) unsigned char scratch[BN_SCRATCH_BYTES];

static pthread_mutex_t scratch_mu = PTHREAD_MUTEX_INITIALIZER;
static void *scratch_addr[BN_SCRATCH_MAX_DEVICES];

extern "C" int bn_scratch_ptr(void **ptr, size_t *bytes) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return BN_ERR_HIP;
  if (dev < 0 || dev >= BN_SCRATCH_MAX_DEVICES) return BN_ERR_ARG;
  pthread_mutex_lock(&scratch_mu);
  if (!scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pthread_mutex_unlock(&scratch_mu);
      return BN_ERR_HIP;
    }
    scratch_addr[dev] = p;
  }
  *ptr = scratch_addr[dev];
  pthread_mutex_unlock(&scratch_mu);
  if (bytes) *bytes = BN_SCRATCH_BYTES;
  return 0;
}
