// Kernel-side helpers shared by the megakernel, replay, wavefront and adjoint
// kernels: block shape, the LDS short-stack block, wave reductions and the
// counter flush.
#pragma once
#include <hip/hip_runtime.h>

#include "mrad_device.hpp"

namespace mrad {

constexpr int kBlock = 256;          // 4 wave64s
constexpr int kWaves = kBlock / 64;

// Per-block LDS for the short traversal stacks (kShort entries per lane) of
// its kW waves.
template <int kShort = kShortStack, int kW = kWaves>
struct StackLdsT {
#if MRAD_STACK_PACKED
	unsigned long long slot[kW * (kShort > 0 ? kShort : 1) * 64];
#else
	int node[kW * (kShort > 0 ? kShort : 1) * 64];
	float entry[kW * (kShort > 0 ? kShort : 1) * 64];
#endif
};
using StackLds = StackLdsT<>;
template <int kShort, int kW>
__device__ __forceinline__ LdsStack laneStack(StackLdsT<kShort, kW> &s) {
	const int base = (threadIdx.x >> 6) * kShort * 64 + (threadIdx.x & 63);
#if MRAD_STACK_PACKED
	return {(LdsSlot *) (s.slot + base)};
#else
	return {(LdsInt *) (s.node + base), (LdsFloat *) (s.entry + base)};
#endif
}

__device__ __forceinline__ unsigned long long waveSum(unsigned long long v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}
__device__ __forceinline__ unsigned long long waveMax(unsigned long long v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		unsigned long long o = __shfl_xor(v, off, 64);
		v = o > v ? o : v;
	}
	return v;
}

// Counters of one wave folded into the 14-slot frame block: one atomic per
// slot per wave instead of the reference's per-thread atomics
// (GpuKernelSources.java:1295-1310).  Every lane of the wave must call it.
template <bool kMetrics>
__device__ __forceinline__ void flushCounters(const Counters<kMetrics> &c, unsigned long long *stats) {
	if constexpr (kMetrics) {
		const int lane = threadIdx.x & 63;
#pragma unroll
		for (int i = 0; i < kStatCount; i++) {
			unsigned long long v = i == 9 ? waveMax(c.v[i]) : waveSum(c.v[i]);
			if (lane == 0) {
				if (i == 9) atomicMax(stats + i, v);
				else atomicAdd(stats + i, v);
			}
		}
	}
}

}  // namespace mrad
