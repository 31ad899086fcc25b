// Kernel-argument prefetch for latency-bound kernels.
#pragma once

namespace avbd {

// Makes every listed value (kernel arguments, blockDim / gridDim) live in
// scalar registers at this point, so that their loads from the argument
// segment are issued together and waited for once, instead of each where it
// is first used, one dependent scalar round trip after another.  Emits no
// instruction of its own.
template <class T>
__device__ __forceinline__ void kernarg_live(T a) {
    asm volatile("" ::"s"(a));
}
template <class... T>
__device__ __forceinline__ void kernarg_prefetch(T... a) {
    (kernarg_live(a), ...);
}

}  // namespace avbd
