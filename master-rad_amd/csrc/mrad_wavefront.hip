// Batched wavefront form of the path tracer (no counterpart in the
// reference, whose renderer is a per-pixel megakernel,
// gpu/GpuKernelSources.java:1167-1315).
//
// Many scenes of uneven size and resolution share one arena and one bounded
// pool of in-flight paths.  Each bounce step is three persistent kernels:
//   generate  fills the pool up to its capacity with new camera paths, taken
//             in global path order across all scenes;
//   trace     closest hit of every live path (the megakernel's traversal);
//   shade     emission, scattering and roulette of every live path; the
//             survivors are compacted into the other pool by a wave64 ballot
//             and prefix sum with one atomic per wave, finished paths write
//             their radiance to a per-path slot.
// Two bounce steps (pool A to B and back) are captured once as a hipGraph and
// replayed until every path has finished.  Every path runs exactly the
// megakernel's arithmetic and every pixel sums its samples in the megakernel's
// order, so both forms produce bit-identical images.
//
// Paths are numbered in global order (scene, pixel, sample) with 64-bit ids,
// and the batch is cut into chunks of whole pixels.  A finished path writes
// its radiance to a slot of one of two result buffers (chunk c uses buffer
// c mod 2); once every path of a chunk has finished, a resolve kernel sums
// each of its pixels' samples in sample order and the buffer is free for chunk
// c + 2.  Chunk c + 2 is generated only after chunk c is resolved, while chunk
// c + 1 keeps the pool full, so result memory is bounded by two chunks (a
// multiple of the pool), not by the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include <string>
#include <vector>

#include "mrad_host.hpp"
#include "mrad_kernels_common.hpp"
#include "mrad_tracer.hpp"

namespace mrad {

struct WfScene {
	SceneRange R;
	float cam[12];
	int width, height, spp, maxDepth, rrStart, frameIndex;
	uint64_t frameSeed;
	uint64_t pathBase;      // first global path id of the scene
	uint64_t pixelBase;     // first global pixel of the scene
};

// A run of whole pixels and their paths; its radiance goes to result buffer
// (index mod 2).
struct WfChunk {
	uint64_t pathStart, pixelStart;
	uint32_t pathCount, pixelCount;
};

// Path pool, structure of arrays of 16-byte records:
//   ori = (o.x, o.y, o.z, 0)     dir  = (d.x, d.y, d.z, T.x)
//   tl  = (T.y, T.z, L.x, L.y)   meta = (rng, result slot, scene | depth << 16, L.z bits)
// The result slot is buffer * chunkPaths + the path's index in its chunk.
// Trace reads ori, dir and meta; shade reads dir and meta, and tl only past
// depth 0 (a camera path starts with T = 1 and L = 0, which generate does not
// store); nothing but trace reads the origin.  Generate stores no origin
// either: every camera ray starts at its scene's camera position
// (poolOrigin).
struct WfPool {
	float4 *ori, *dir, *tl;
	uint4 *meta;
};

// Device counter block: pool counts, the chunk being generated and how many
// of its paths are generated, how many chunks may be generated (raised by the
// host as chunks are resolved), the trace kernel's work counter, and the
// finished paths of the chunk in each result buffer.
enum {
	kCount0 = 0, kCount1 = 1, kGenChunk = 2, kGenOffset = 3, kGenLimit = 4, kWorkTrace = 5, kDone0 = 6, kDone1 = 7,
	kSuspend0 = 8, kSuspend1 = 9, kWorkPass1 = 10, kWorkPass2 = 11,
	kCounterWords = 16
};

// Continuation list of the capped trace passes: per suspended closest-hit
// query its pool slot, best distance bits, best rank and stack size (hdr),
// and its stack entries, entry j of item k at stack[j * capacity + k].
struct WfCont {
	uint4 *hdr;
	unsigned long long *stack;
};
constexpr uint32_t kDead = 0xFFFFu;          // depth field of a path that never traces
constexpr uint32_t kMissKind = 7u;

struct WfParams {
	DevScene S;
	const WfScene *scenes;
	int sceneCount;
	const WfChunk *chunks;
	uint32_t chunkCount;
	uint32_t chunkPaths;    // result slots per buffer
	WfPool pool[2];
	float4 *hitA;           // resolved closest hit: position, material index bits
	float4 *hitB;           //                       normal, 1 if hit else 0
	float *results;         // 3 floats per result slot, two buffers of chunkPaths slots
	float *images;          // 3 floats per pixel, all scenes
	uint32_t *counters;
	uint32_t capacity;
	unsigned long long *stats;
	WfCont cont[2];         // continuation lists of the capped trace passes
	uint32_t contCap;       // items per list
};

__device__ __forceinline__ int sceneOfPath(const WfParams &P, uint64_t g) {
	int lo = 0, hi = P.sceneCount - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (P.scenes[mid].pathBase <= g) lo = mid; else hi = mid - 1;
	}
	return lo;
}
__device__ __forceinline__ int sceneOfPixel(const WfParams &P, uint64_t p) {
	int lo = 0, hi = P.sceneCount - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (P.scenes[mid].pixelBase <= p) lo = mid; else hi = mid - 1;
	}
	return lo;
}

// Order in which a scene's pixels are given path ids (each pixel's samples
// consecutive).  1: 8x8 pixel tiles in raster order, Z-order inside a full
// tile and row order inside the partial tiles of the right and bottom edges,
// so the 64 paths of a trace wave cover a compact patch of the image (a 4x4
// patch at 4 samples, 2x2 at 16) instead of a piece of one row; 0: row
// order.  Paths of a pixel, their seeds and the per-pixel sums are the same
// either way, so images are too.
#ifndef MRAD_WF_PIXEL_TILES
#define MRAD_WF_PIXEL_TILES 0
#endif

// The pixel of rank r of a width x height image in that order.
__device__ __forceinline__ void rankToPixel(uint32_t r, int width, int height, int &x, int &y) {
	if (!MRAD_WF_PIXEL_TILES) {
		x = (int) (r % (uint32_t) width);
		y = (int) (r / (uint32_t) width);
		return;
	}
	const uint32_t W = (uint32_t) width, H = (uint32_t) height;
	const uint32_t fullBands = H / 8u, bandPixels = 8u * W;
	uint32_t band, off, bh;
	if (r < fullBands * bandPixels) {
		band = r / bandPixels;
		off = r - band * bandPixels;
		bh = 8u;
	} else {
		band = fullBands;
		off = r - fullBands * bandPixels;
		bh = H - 8u * fullBands;
	}
	const uint32_t fullBlocks = W / 8u, blockPixels = 8u * bh;
	uint32_t block, in, bw;
	if (off < fullBlocks * blockPixels) {
		block = off / blockPixels;
		in = off - block * blockPixels;
		bw = 8u;
	} else {
		block = fullBlocks;
		in = off - fullBlocks * blockPixels;
		bw = W - 8u * fullBlocks;
	}
	uint32_t lx, ly;
	if (bw == 8u && bh == 8u) {
		lx = (in & 1u) | ((in >> 1) & 2u) | ((in >> 2) & 4u);
		ly = ((in >> 1) & 1u) | ((in >> 2) & 2u) | ((in >> 3) & 4u);
	} else {
		lx = in % bw;
		ly = in / bw;
	}
	x = (int) (8u * block + lx);
	y = (int) (8u * band + ly);
}

// The origin of pool entry i: a camera path's (depth 0) is its scene's camera
// position, cameraRay's origin, which generate does not store.
__device__ __forceinline__ float4 poolOrigin(const WfParams &P, const WfPool &Q, uint32_t i, const uint4 &meta) {
	if ((meta.z >> 16) == 0u) {
		const float *c = P.scenes[meta.z & 0xFFFFu].cam;
		return make_float4(c[0], c[1], c[2], 0.0f);
	}
	return Q.ori[i];
}

// Paths the next generate step appends to pool `cur`: up to the pool's free
// space, from the current chunk only, and none from a chunk the host has not
// released yet.
__device__ __forceinline__ uint32_t genCount(const WfParams &P, int cur) {
	const uint32_t chunk = P.counters[kGenChunk];
	if (chunk >= P.chunkCount || chunk >= P.counters[kGenLimit]) return 0u;
	const uint32_t base = P.counters[kCount0 + cur];
	const uint32_t free = P.capacity > base ? P.capacity - base : 0u;
	const uint32_t left = P.chunks[chunk].pathCount - P.counters[kGenOffset];
	return free < left ? free : left;
}

// Camera paths of the current chunk appended to pool `cur` (renderKernel's
// seed and camera ray, GpuKernelSources.java:1231-1264).
__global__ __launch_bounds__(kBlock) void wfGenerate(WfParams P, int cur) {
	const uint32_t n = genCount(P, cur);
	if (n == 0) return;
	const uint32_t base = P.counters[kCount0 + cur];
	const uint32_t chunk = P.counters[kGenChunk];
	const uint32_t offset = P.counters[kGenOffset];
	const uint64_t first = P.chunks[chunk].pathStart + offset;
	const uint32_t slotBase = (chunk & 1u) * P.chunkPaths + offset;
	const WfPool Q = P.pool[cur];
	for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
		const uint64_t g = first + j;
		const int k = sceneOfPath(P, g);
		const WfScene &V = P.scenes[k];
		const uint64_t local = g - V.pathBase;
		const uint32_t rank = (uint32_t) (local / (uint64_t) V.spp), s = (uint32_t) (local % (uint64_t) V.spp);
		int x, y;
		rankToPixel(rank, V.width, V.height, x, y);
		const uint32_t pixel = (uint32_t) y * (uint32_t) V.width + (uint32_t) x;
		uint32_t rng = path_seed(V.frameSeed, V.frameIndex, pixel, (int) s);
		f3 o, d;
		cameraRay(V.cam, V.width, V.height, x, y, rng, o, d);
		const uint32_t result = slotBase + j;
		uint32_t depth = 0;
		if (V.maxDepth == 0) {
			depth = kDead;
			P.results[3 * (size_t) result] = 0.0f;
			P.results[3 * (size_t) result + 1] = 0.0f;
			P.results[3 * (size_t) result + 2] = 0.0f;
		}
		const uint32_t slot = base + j;
		Q.dir[slot] = make_float4(d.x, d.y, d.z, 1.0f);          // o is the camera position (poolOrigin)
		Q.meta[slot] = make_uint4(rng, result, (uint32_t) k | (depth << 16), __float_as_uint(0.0f));
	}
}

__global__ void wfAdvanceGenerate(WfParams P, int cur) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	const uint32_t n = genCount(P, cur);
	P.counters[kCount0 + cur] += n;
	if (n > 0) {
		const uint32_t chunk = P.counters[kGenChunk];
		const uint32_t offset = P.counters[kGenOffset] + n;
		const bool done = offset == P.chunks[chunk].pathCount;
		P.counters[kGenOffset] = done ? 0u : offset;
		P.counters[kGenChunk] = done ? chunk + 1u : chunk;
	}
	P.counters[kWorkTrace] = 0;
	P.counters[kWorkPass1] = 0;
	P.counters[kWorkPass2] = 0;
	P.counters[kSuspend0] = 0;
	P.counters[kSuspend1] = 0;
}

// Counter block at the start of a render: zero, with chunks 0 and 1 released.
__global__ void wfInitCounters(WfParams P) {
	if (blockIdx.x != 0 || threadIdx.x >= kCounterWords) return;
	P.counters[threadIdx.x] = threadIdx.x == kGenLimit ? 2u : 0u;
}

// Chunks [first, last) have been resolved: their buffers' finished counts
// start again from zero and chunks up to last + 1 may be generated.  The
// counters are updated by this kernel rather than by host copies, so every
// write to them is ordered with the kernels' atomics on the device.
__global__ void wfRelease(WfParams P, uint32_t first, uint32_t last) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	for (uint32_t c = first; c < last && c < first + 2; c++) P.counters[kDone0 + (c & 1u)] = 0u;
	P.counters[kGenLimit] = last + 2u;
}

__global__ void wfAdvanceShade(WfParams P, int cur) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	P.counters[kCount0 + cur] = 0;
}

// The trace kernels resolve the winner (position, normal, material) with the
// accepting routine's arithmetic, so shading never fetches primitive records.
__device__ __forceinline__ void writeHit(const WfParams &P, uint32_t i, f3 o, f3 d, const Best &best, bool found) {
	if (found) {
		const HitDetail h = resolveHit(P.S, o, d, best.t, best.kind, best.index);
		P.hitA[i] = make_float4(h.pos.x, h.pos.y, h.pos.z, __int_as_float(h.matIndex));
		P.hitB[i] = make_float4(h.normal.x, h.normal.y, h.normal.z, 1.0f);
	} else {
		P.hitB[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
}

// Next chunk of 64 items for this wave from a work counter.
__device__ __forceinline__ uint32_t fetchChunk(uint32_t *counter) {
	uint32_t base = 0;
	if ((threadIdx.x & 63) == 0) base = atomicAdd(counter, 64u);
	return __shfl(base, 0, 64);
}

// Work distribution of the trace kernel.  The waves of one block share a CU,
// and so its L1: a block takes the pool in super-chunks of kTraceSuper
// consecutive paths (neighbouring pixels, or paths that were neighbours a
// bounce earlier), its waves taking 64-path chunks of the current one in
// turn from a block counter in LDS, so that their node fetches hit the same
// lines.  The first MRAD_TRACE_STATIC_PCT % of the pool (whole rounds of
// super-chunks) is dealt out statically and interleaved: block b takes
// super-chunks b, b + B, b + 2B, ... of the B blocks, so every block's share
// samples the whole pool and the shares cost about the same.  The rest is a
// global queue of 64-path chunks that evens out what is left.  kTraceSuper 0
// keeps the single global queue.
#ifndef MRAD_TRACE_SUPER
#define MRAD_TRACE_SUPER 2048
#endif
#ifndef MRAD_TRACE_STATIC_PCT
#define MRAD_TRACE_STATIC_PCT 60
#endif
constexpr uint32_t kTraceSuper = MRAD_TRACE_SUPER;
static_assert(kTraceSuper % 64 == 0, "super-chunks hold whole 64-path chunks");

struct TraceWork {
	uint32_t perBlock;    // paths of the static part per block (whole super-chunks)
	uint32_t staticEnd;   // end of the static part: perBlock * gridDim.x
};
__device__ __forceinline__ TraceWork traceWork(uint32_t count) {
	TraceWork w{0u, 0u};
	if constexpr (kTraceSuper > 0) {
		const uint64_t round = (uint64_t) gridDim.x * kTraceSuper;
		const uint64_t rounds = (uint64_t) count * MRAD_TRACE_STATIC_PCT / 100u / round;
		w.perBlock = (uint32_t) (rounds * kTraceSuper);
		w.staticEnd = (uint32_t) (rounds * round);
	}
	return w;
}
// Next 64-path chunk of this wave: from the block's static share while it
// lasts (`next` is the block's LDS counter), then from the global queue.
__device__ __forceinline__ uint32_t fetchTraceChunk(uint32_t *next, uint32_t *queue, const TraceWork &w) {
	if constexpr (kTraceSuper == 0) {
		return fetchChunk(queue);
	} else {
		uint32_t base = 0;
		if ((threadIdx.x & 63) == 0) {
			const uint32_t off = atomicAdd(next, 64u);
			if (off < w.perBlock)
				base = ((off / kTraceSuper) * gridDim.x + blockIdx.x) * kTraceSuper + off % kTraceSuper;
			else
				base = w.staticEnd + atomicAdd(queue, 64u);
		}
		return __shfl(base, 0, 64);
	}
}

// Threads per block of the trace kernels (wfTrace, wfTracePass): the waves
// of a block share its super-chunks (and its CU's L1).
#ifndef MRAD_TRACE_BLOCK
#define MRAD_TRACE_BLOCK 256
#endif
constexpr int kTraceBlock = MRAD_TRACE_BLOCK;
constexpr int kTraceWaves = kTraceBlock / 64;
static_assert(kTraceBlock % 64 == 0 && kTraceBlock <= 1024, "trace blocks hold whole waves");

// Occupancy the trace kernels are compiled for (waves per SIMD; the register
// budget is 512 / this).  7 waves (72 VGPRs) with the packed stack slots and
// the branch-free push is 1.8 % faster than 8 waves (64 VGPRs), where both
// spill live values inside the traversal loop.
#ifndef MRAD_TRACE_WAVES_PER_EU
#define MRAD_TRACE_WAVES_PER_EU 7
#endif

// Node slots of a scene's top BVH levels each trace wave keeps in LDS (0:
// none).  A wave reloads them when the scene of its chunk's first path
// changes; paths of other scenes in the chunk read every node from HBM.
#ifndef MRAD_TRACE_TOP
#define MRAD_TRACE_TOP 0
#endif
constexpr int kTraceTop = MRAD_TRACE_TOP;

// Octant of a direction: bit 0 for x < 0, bit 1 for y < 0, bit 2 for z < 0.
__device__ __forceinline__ int octantOf(float dx, float dy, float dz) {
	return (dx < 0.0f ? 1 : 0) | (dy < 0.0f ? 2 : 0) | (dz < 0.0f ? 4 : 0);
}

// Closest hit of every live path in pool `cur` (persistent).  With kFuse the
// same kernel also shades each path once the wave's 64 traversals have ended
// (all lanes active, the traversal state dead) and appends the survivors to
// pool cur ^ 1, grouped by the octant of their new direction within the wave,
// with one atomic per wave on the destination count; finished paths write
// their radiance and are counted per result buffer as wfShade counts them.
// The hit records are then never written.
template <bool kMetrics, int kCap, int kShort, bool kFuse = false>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(MRAD_TRACE_WAVES_PER_EU)))
void wfTrace(WfParams P, int cur) {
	__shared__ StackLdsT<kShort, kTraceWaves> stackLds;
	const LdsStack lds = laneStack(stackLds);
	constexpr int kTop = kTraceTop;
	__shared__ float4 topLds[kTraceWaves][kTop > 0 ? 4 * kTop : 1];
	__shared__ uint32_t doneLds[2];
	__shared__ uint32_t nextLds;
	const uint32_t count = P.counters[kCount0 + cur];
	const WfPool Q = P.pool[cur];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned long long below = (1ull << lane) - 1ull;
	uint32_t finished[2] = {0u, 0u};
	if (kFuse && threadIdx.x < 2) doneLds[threadIdx.x] = 0u;
	const TraceWork work = traceWork(count);
	if constexpr (kTraceSuper > 0) {
		if (threadIdx.x == 0) nextLds = 0u;
		__syncthreads();
	}
	int cached = -1;
	TopNodes top{(const LdsFloat4 *) topLds[wave], 0, 0};
	Counters<kMetrics> c;
	c.zero();
	for (;;) {
		const uint32_t base = fetchTraceChunk(&nextLds, P.counters + kWorkTrace, work);
		if (base >= count) break;
		const uint32_t i = base + lane;
		uint4 meta = make_uint4(0, 0, 0xFFFFFFFFu, 0);
		if (i < count) meta = Q.meta[i];
		if constexpr (kTop > 0) {
			// lane 0's path (base < count) names the scene staged for the chunk
			const int lead = __shfl((int) (meta.z & 0xFFFFu), 0, 64);
			if (lead != cached) {
				const SceneRange &L = P.scenes[lead].R;
				const int n = L.useBvh ? min(kTop, L.topSlots) : 0;
				for (int k = lane; k < 4 * n; k += 64) topLds[wave][k] = P.S.nodes[4 * (size_t) L.nodeBase + k];
				__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
				__builtin_amdgcn_wave_barrier();
				cached = lead;
				top.base = L.nodeBase;
				top.count = n;
			}
		}
		if constexpr (!kFuse) {
			if (i < count) {
				if ((meta.z >> 16) == kDead) {
					P.hitB[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
					continue;
				}
				const float4 a = poolOrigin(P, Q, i, meta), b = Q.dir[i];
				const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
				const SceneRange &R = P.scenes[meta.z & 0xFFFFu].R;
				c.add(0, 1);
				Best best;
				const bool found = traceClosest<kMetrics, kCap, kShort, 0, kTop>(P.S, R, o, d, best, c, lds, top);
				writeHit(P, i, o, d, best, found);
			}
		} else {
			bool alive = false;
			float4 n0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), n1 = n0, n2 = n0;
			if (i < count) {
				const uint32_t depth = meta.z >> 16;
				const uint32_t buffer = meta.y >= P.chunkPaths ? 1u : 0u;
				if (depth == kDead) {
					finished[buffer]++;
				} else {
					const float4 a = poolOrigin(P, Q, i, meta), b = Q.dir[i];
					f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
					const WfScene &V = P.scenes[meta.z & 0xFFFFu];
					c.add(0, 1);
					Best best;
					const bool found =
					    traceClosest<kMetrics, kCap, kShort, 0, kTop>(P.S, V.R, o, d, best, c, lds, top);
					HitDetail h{};
					if (found) h = resolveHit(P.S, o, d, best.t, best.kind, best.index);
					const float4 t4 = depth == 0 ? make_float4(1.0f, 1.0f, 0.0f, 0.0f) : Q.tl[i];
					f3 T = mk(b.w, t4.x, t4.y), L = mk(t4.z, t4.w, __uint_as_float(meta.w));
					uint32_t rng = meta.x;
					alive = shadeResolved(P.S, V.R.bg, found, h, (int) depth, V.maxDepth, V.rrStart, o, d, T, L,
					                      rng) &&
					        (int) depth + 1 < V.maxDepth;
					if (alive) {
						n0 = make_float4(o.x, o.y, o.z, 0.0f);
						n1 = make_float4(d.x, d.y, d.z, T.x);
						n2 = make_float4(T.y, T.z, L.x, L.y);
						meta.x = rng;
						meta.z = (meta.z & 0xFFFFu) | ((depth + 1) << 16);
						meta.w = __float_as_uint(L.z);
					} else {
						float *r = P.results + 3 * (size_t) meta.y;
						r[0] = L.x;
						r[1] = L.y;
						r[2] = L.z;
						finished[buffer]++;
					}
				}
			}
			// survivors of the wave, stable within each octant of the new direction
			const int bucket = alive ? octantOf(n1.x, n1.y, n1.z) : 0;
			uint32_t before = 0, total = 0;
			unsigned long long mine = 0ull;
			for (int q = 0; q < 8; q++) {
				const unsigned long long m = __ballot(alive && bucket == q);
				if (q == bucket) {
					mine = m;
					before = total;
				}
				total += (uint32_t) __popcll(m);
			}
			uint32_t dest = 0;
			if (lane == 0 && total) dest = atomicAdd(P.counters + kCount0 + (cur ^ 1), total);
			dest = __shfl(dest, 0, 64);
			if (alive) {
				const WfPool D = P.pool[cur ^ 1];
				const uint32_t slot = dest + before + (uint32_t) __popcll(mine & below);
				D.ori[slot] = n0;
				D.dir[slot] = n1;
				D.tl[slot] = n2;
				D.meta[slot] = meta;
			}
		}
	}
	if (P.stats) flushCounters<kMetrics>(c, P.stats);
	if constexpr (kFuse) {
		// finished paths per result buffer: one LDS sum per block, one atomic
		// per block and buffer
		__syncthreads();
		for (int b = 0; b < 2; b++) {
			uint32_t v = finished[b];
			for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
			if (lane == 0 && v) atomicAdd(&doneLds[b], v);
		}
		__syncthreads();
		if (threadIdx.x < 2 && doneLds[threadIdx.x]) atomicAdd(P.counters + kDone0 + threadIdx.x, doneLds[threadIdx.x]);
	}
}

// ------------------------------------------------------ capped trace passes
//
// The trace of a bounce step in passes.  A wave of the single pass runs until
// its longest traversal ends, and traversal lengths are heavy-tailed (the
// batch's secondary rays: ~55 pops on average, up to ~300), so most of its
// lanes idle while a few finish.  Here pass 0 starts the closest-hit query of
// every live path of the pool and suspends any query still traversing after
// `budget` pops: its pool slot, best distance and rank, stack size and stack
// entries go to a continuation list.  The next pass resumes the suspended
// queries, 64 to a wave, from that list; the last pass has no budget.  A
// resumed query makes the same pops in the same order from the same state, so
// hits, images and counters are those of the single pass.  A query whose
// stack holds more than kSuspendEntries entries, or that finds its list full,
// is not suspended: it keeps traversing in the current pass.
#ifndef MRAD_PASS_CAMERA_UNCAPPED
#define MRAD_PASS_CAMERA_UNCAPPED 1
#endif
// A wave that still has `dense` or more lanes traversing when its budget runs
// out is not split: it keeps going for MRAD_PASS_RECHECK_POPS more pops and is
// checked again, so only waves that have thinned out hand their queries to the
// next pass (dense <= 0: every wave is suspended at the budget).
#ifndef MRAD_PASS_RECHECK_POPS
#define MRAD_PASS_RECHECK_POPS 32
#endif
constexpr int kPassRecheckPops = MRAD_PASS_RECHECK_POPS;
static_assert(kPassRecheckPops >= 1, "dense-wave recheck interval");
#ifndef MRAD_SUSPEND_ENTRIES
#define MRAD_SUSPEND_ENTRIES 12
#endif
constexpr int kSuspendEntries = MRAD_SUSPEND_ENTRIES;
static_assert(kSuspendEntries >= 1 && kSuspendEntries <= 32, "suspended stacks fit the smallest storage");

template <bool kMetrics, int kCap, int kShort>
__device__ __forceinline__ void bvhLoopAnyForm(const DevScene &S, const SceneRange &R, const Ray &ray, Best &best,
                                               Counters<kMetrics> &c, BvhState<kCap, kShort> &b, int budget) {
	// one loop per slab-quotient form, as traceClosest chooses them
	if (ray.fast && ray.d.x != 0.0f && ray.d.y != 0.0f && ray.d.z != 0.0f)
		bvhLoop<kMetrics, kCap, kShort, kSlabFastNz, 0, true>(S, R, ray, best, c, b, TopNodes{}, budget);
	else if (ray.fast)
		bvhLoop<kMetrics, kCap, kShort, kSlabFast, 0, true>(S, R, ray, best, c, b, TopNodes{}, budget);
	else
		bvhLoop<kMetrics, kCap, kShort, kSlabIeee, 0, true>(S, R, ray, best, c, b, TopNodes{}, budget);
}

// Pass `pass` (0, 1 or 2) of a bounce step on pool `cur`; budget <= 0 for the
// last pass.  Pass p > 0 reads list (p - 1) & 1; pass p writes list p & 1.
// `dense`: the dense-wave threshold above.
template <bool kMetrics, int kCap, int kShort>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(MRAD_TRACE_WAVES_PER_EU)))
void wfTracePass(WfParams P, int cur, int pass, int budget, int dense) {
	__shared__ StackLdsT<kShort, kTraceWaves> stackLds;
	const LdsStack lds = laneStack(stackLds);
	const int lane = threadIdx.x & 63;
	const unsigned long long below = (1ull << lane) - 1ull;
	const WfPool Q = P.pool[cur];
	const WfCont in = P.cont[(pass + 1) & 1], out = P.cont[pass & 1];
	// a list's count may exceed its capacity: items past it were not stored
	// (their queries finished in the pass that tried to suspend them)
	const uint32_t count = pass == 0 ? P.counters[kCount0 + cur]
	                                 : min(P.counters[kSuspend0 + ((pass + 1) & 1)], P.contCap);
	uint32_t *work = P.counters + (pass == 0 ? kWorkTrace : pass == 1 ? kWorkPass1 : kWorkPass2);
	uint32_t *outCount = P.counters + kSuspend0 + (pass & 1);
	// pass 0 reads the pool in super-chunks, as wfTrace does; a continuation
	// list has no order worth keeping
	__shared__ uint32_t nextLds;
	const TraceWork plan = pass == 0 ? traceWork(count) : TraceWork{0u, 0u};
	if constexpr (kTraceSuper > 0) {
		if (threadIdx.x == 0) nextLds = 0u;
		__syncthreads();
	}
	Counters<kMetrics> c;
	c.zero();
	for (;;) {
		const uint32_t base = fetchTraceChunk(&nextLds, work, plan);
		if (base >= count) break;
		const uint32_t k = base + lane;
		bool live = k < count;
		uint32_t i = k;                   // pool slot of the path
		uint4 hdr = make_uint4(0u, 0u, 0u, 0u);
		if (live && pass > 0) {
			hdr = in.hdr[k];
			i = hdr.x;
		}
		uint4 meta = make_uint4(0u, 0u, 0u, 0u);
		if (live) meta = Q.meta[i];
		if (live && pass == 0 && (meta.z >> 16) == kDead) {
			P.hitB[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			live = false;
		}
		const SceneRange &R = P.scenes[live ? (int) (meta.z & 0xFFFFu) : 0].R;
		Best best{1e30f, 0xFFFFFFFFu, -1, -1};
		BvhState<kCap, kShort> b;
		b.st.lds = lds;
		b.size = 0;
		b.overflow = false;
		Ray ray{};
		bool traversing = false;
		if (live) {
			const float4 a4 = poolOrigin(P, Q, i, meta), d4 = Q.dir[i];
			ray = makeRay(mk(a4.x, a4.y, a4.z), mk(d4.x, d4.y, d4.z), R.fastSlabs != 0);
			if (pass == 0) {
				c.add(0, 1);
				if (R.useBvh != 0) bvhBegin<kMetrics, kCap, kShort, kSlabAny>(R, ray, best, c, b);
			} else {
				// only BVH queries are suspended; rank 0xFFFFFFFF is "no hit yet"
				// (arena indices stay below kRefIndexMask)
				best.t = __uint_as_float(hdr.y);
				best.rank = hdr.z;
				best.kind = hdr.z == 0xFFFFFFFFu ? -1 : (int) (hdr.z >> 30);
				best.index = hdr.z == 0xFFFFFFFFu ? -1 : (int) (hdr.z & kRefIndexMask);
				b.size = (int) hdr.w;
				for (int j = 0; j < b.size; j++) {
					const unsigned long long v = in.stack[(size_t) j * P.contCap + k];
					b.st.set(j, (int) (uint32_t) v, __uint_as_float((uint32_t) (v >> 32)));
				}
			}
			traversing = R.useBvh != 0;
		}
		bool suspended = false;
		int left = budget > 0 ? budget : 0x7FFFFFFF;
#if MRAD_PASS_CAMERA_UNCAPPED
		// a wave of camera paths only is not capped: camera rays of a wave
		// (neighbouring pixels, or one pixel's samples) traverse alike, and
		// suspending them costs their locality for little regrouping
		if (pass == 0 && __ballot(live && (meta.z >> 16) != 0u) == 0ull) left = 0x7FFFFFFF;
#endif
		for (;;) {
			if (traversing) bvhLoopAnyForm<kMetrics, kCap, kShort>(P.S, R, ray, best, c, b, left);
			const bool open = traversing && !b.overflow && b.size > 0;
			if (dense > 0 && left != 0x7FFFFFFF && __popcll(__ballot(open)) >= dense) {
				left = kPassRecheckPops;   // still dense: every open lane pops again before the next check
				continue;
			}
			const bool want = open && b.size <= kSuspendEntries;
			const unsigned long long m = __ballot(want);
			uint32_t first = 0;
			if (m) {
				const int leader = __ffsll((long long) m) - 1;
				if (lane == leader) first = atomicAdd(outCount, (uint32_t) __popcll(m));
				first = __shfl(first, leader, 64);
			}
			const uint32_t slot = first + (uint32_t) __popcll(m & below);
			if (want && slot < P.contCap) {
				out.hdr[slot] = make_uint4(i, __float_as_uint(best.t), best.rank, (uint32_t) b.size);
				for (int j = 0; j < b.size; j++) {
					int n;
					float e;
					b.st.get(j, n, e);
					out.stack[(size_t) j * P.contCap + slot] =
					    (uint32_t) n | (unsigned long long) __float_as_uint(e) << 32;
				}
				suspended = true;
			}
			if (!open || suspended) break;
			left = 0x7FFFFFFF;            // not suspended: finish in this pass
		}
		if (live && !suspended) {
			if (R.useBvh != 0) bvhEnd<kMetrics>(P.S, R, ray, best, c, b);
			else traceLinear<kMetrics>(P.S, R, ray, best, c);
			c.add(2, (uint32_t) R.planeCount);
			c.add(5, (uint32_t) R.planeCount);
			tracePlanes(P.S, R, ray.o, ray.d, best);
			writeHit(P, i, ray.o, ray.d, best, best.kind >= 0);
		}
	}
	if (P.stats) flushCounters<kMetrics>(c, P.stats);
}

// Shading of one pool entry; true when the path continues (its new state is
// left in ori, dir, tl, meta), false when it finished (its radiance is written
// to its result slot) or the entry is empty.  A path that finishes here, or
// that never traced (depth 0), adds one to finished[its result buffer].
__device__ __forceinline__ bool shadeEntry(const WfParams &P, const WfPool &Q, uint32_t i, uint32_t count, float4 &ori,
                                           float4 &dir, float4 &tl, uint4 &meta, uint32_t (&finished)[2]) {
	if (i >= count) return false;
	meta = Q.meta[i];
	const uint32_t depth = meta.z >> 16;
	const uint32_t buffer = meta.y >= P.chunkPaths ? 1u : 0u;
	if (depth == kDead) {
		finished[buffer]++;
		return false;
	}
	dir = Q.dir[i];
	tl = depth == 0 ? make_float4(1.0f, 1.0f, 0.0f, 0.0f) : Q.tl[i];
	f3 o, d = mk(dir.x, dir.y, dir.z);          // o: the new origin (the incoming one is not needed)
	f3 T = mk(dir.w, tl.x, tl.y), L = mk(tl.z, tl.w, __uint_as_float(meta.w));
	uint32_t rng = meta.x;
	const WfScene &V = P.scenes[meta.z & 0xFFFFu];
	const float4 ha = P.hitA[i], hb = P.hitB[i];
	const bool found = hb.w != 0.0f;
	HitDetail h;
	h.pos = mk(ha.x, ha.y, ha.z);
	h.matIndex = __float_as_int(ha.w);
	h.normal = mk(hb.x, hb.y, hb.z);
	const bool alive = shadeResolved(P.S, V.R.bg, found, h, (int) depth, V.maxDepth, V.rrStart, o, d, T, L, rng) &&
	                   (int) depth + 1 < V.maxDepth;
	if (alive) {
		ori = make_float4(o.x, o.y, o.z, 0.0f);
		dir = make_float4(d.x, d.y, d.z, T.x);
		tl = make_float4(T.y, T.z, L.x, L.y);
		meta.x = rng;
		meta.z = (meta.z & 0xFFFFu) | ((depth + 1) << 16);
		meta.w = __float_as_uint(L.z);
	} else {
		float *r = P.results + 3 * (size_t) meta.y;
		r[0] = L.x;
		r[1] = L.y;
		r[2] = L.z;
		finished[buffer]++;
	}
	return alive;
}

// Shading and compaction of every live path of pool `cur` into pool cur ^ 1.
// Shading costs the same for every path, so blocks take fixed tiles of
// kShadeTile entries (kShadePer per thread) in turn; the survivors of a tile
// keep their order and take one atomic per tile on the destination count
// (per-wave atomics on one address serialise and cost more than the shading).
#ifndef MRAD_SHADE_PER
#define MRAD_SHADE_PER 4
#endif
constexpr int kShadePer = MRAD_SHADE_PER;
constexpr uint32_t kShadeTile = kBlock * kShadePer;

#ifndef MRAD_SHADE_OCTANTS
#define MRAD_SHADE_OCTANTS 1
#endif

__global__ __launch_bounds__(kBlock) void wfShade(WfParams P, int cur) {
#if MRAD_SHADE_OCTANTS
	// Survivors of a tile are grouped by the octant of their new direction
	// (a counting sort over 8 buckets, stable within a bucket), so a trace
	// wave's 64 consecutive paths start near each other and head the same way.
	constexpr int kBuckets = 8;
#else
	constexpr int kBuckets = 1;
#endif
	__shared__ uint32_t part[kBuckets * kShadePer * kWaves];
	__shared__ uint32_t tileBase;
	__shared__ uint32_t doneLds[2];
	const uint32_t count = P.counters[kCount0 + cur];
	const WfPool Q = P.pool[cur], D = P.pool[cur ^ 1];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned long long below = (1ull << lane) - 1ull;
	uint32_t finished[2] = {0u, 0u};
	if (threadIdx.x < 2) doneLds[threadIdx.x] = 0u;
	for (uint32_t tile = blockIdx.x * kShadeTile; tile < count; tile += gridDim.x * kShadeTile) {
		float4 ori[kShadePer], dir[kShadePer], tl[kShadePer];
		uint4 meta[kShadePer];
		bool alive[kShadePer];
		int bucket[kShadePer];
		unsigned long long mask[kShadePer];
#pragma unroll
		for (int k = 0; k < kShadePer; k++) {
			alive[k] = shadeEntry(P, Q, tile + k * kBlock + threadIdx.x, count, ori[k], dir[k], tl[k], meta[k],
			                      finished);
			bucket[k] = (kBuckets == 1 || !alive[k]) ? 0
			          : octantOf(dir[k].x, dir[k].y, dir[k].z);
			for (int q = 0; q < kBuckets; q++) {
				const unsigned long long m = __ballot(alive[k] && bucket[k] == q);
				if (q == bucket[k]) mask[k] = m;
				if (lane == 0) part[(q * kShadePer + k) * kWaves + wave] = (uint32_t) __popcll(m);
			}
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			uint32_t sum = 0;
			for (int j = 0; j < kBuckets * kShadePer * kWaves; j++) {
				const uint32_t n = part[j];
				part[j] = sum;
				sum += n;
			}
			tileBase = sum ? atomicAdd(P.counters + kCount0 + (cur ^ 1), sum) : 0u;
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < kShadePer; k++) {
			if (alive[k]) {
				const uint32_t slot = tileBase + part[(bucket[k] * kShadePer + k) * kWaves + wave] +
				                      (uint32_t) __popcll(mask[k] & below);
				D.ori[slot] = ori[k];
				D.dir[slot] = dir[k];
				D.tl[slot] = tl[k];
				D.meta[slot] = meta[k];
			}
		}
		__syncthreads();
	}
	// finished paths per result buffer: one LDS sum per block, one atomic per
	// block and buffer
	__syncthreads();
	for (int b = 0; b < 2; b++) {
		uint32_t v = finished[b];
		for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
		if (lane == 0 && v) atomicAdd(&doneLds[b], v);
	}
	__syncthreads();
	if (threadIdx.x < 2 && doneLds[threadIdx.x]) atomicAdd(P.counters + kDone0 + threadIdx.x, doneLds[threadIdx.x]);
}

// Sums the samples of every pixel of chunk `c` in sample order.
__global__ __launch_bounds__(kBlock) void wfResolve(WfParams P, uint32_t c) {
	const WfChunk C = P.chunks[c];
	const size_t slotBase = (size_t) (c & 1u) * P.chunkPaths;
	for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < C.pixelCount; j += gridDim.x * kBlock) {
		const uint64_t p = C.pixelStart + j;        // the pixel of this rank in the batch's rank order
		const WfScene &V = P.scenes[sceneOfPixel(P, p)];
		const uint64_t rank = p - V.pixelBase;
		const size_t first = slotBase + (size_t) (V.pathBase + rank * (uint64_t) V.spp - C.pathStart);
		const float *r = P.results + 3 * first;
		f3 acc = mk(0.0f, 0.0f, 0.0f);
		if ((V.spp & 3) == 0 && (first & 3) == 0) {
			// four samples (12 floats) per three 16-byte loads, summed in sample order
			const float4 *q = reinterpret_cast<const float4 *>(r);
			for (int s = 0; s < V.spp; s += 4) {
				const float4 a = q[0], b = q[1], c = q[2];
				q += 3;
				acc = s == 0 ? mk(a.x, a.y, a.z) : acc + mk(a.x, a.y, a.z);
				acc = acc + mk(a.w, b.x, b.y);
				acc = acc + mk(b.z, b.w, c.x);
				acc = acc + mk(c.y, c.z, c.w);
			}
			acc = acc * (1.0f / (float) V.spp);
		} else {
			for (int s = 0; s < V.spp; s++) {
				const f3 col = mk(r[3 * s], r[3 * s + 1], r[3 * s + 2]);
				acc = s == 0 ? col : acc + col;
				if (s == V.spp - 1) acc = acc * (1.0f / (float) V.spp);
			}
		}
		int x, y;
		rankToPixel((uint32_t) rank, V.width, V.height, x, y);
		const size_t out = 3 * ((size_t) V.pixelBase + (size_t) y * (size_t) V.width + (size_t) x);
		P.images[out] = acc.x;
		P.images[out + 1] = acc.y;
		P.images[out + 2] = acc.z;
	}
}

// ------------------------------------------------ primitive-class regrouping
//
// Variant of wfTrace in which the leaf tests of a wave are regrouped by
// primitive class.  Each lane traverses interior nodes until it reaches a leaf
// (or finishes); then the refs of all lanes' current leaves are appended to a
// per-wave task list in LDS, bucketed by class with a count pass and a prefix,
// and the whole wave runs the sphere tasks, then the box tasks, then the
// affine ones, 64 at a time with uniform control flow, so an expensive affine
// test on one lane no longer serialises the cheap tests of the others.
//
// Every ref of a leaf is tested against the lane's best distance at leaf
// entry and the lane's best is the minimum of (t, rank) over the accepted
// candidates (an LDS 64-bit atomic min).  The reference tests a leaf's refs in
// order against the running best (GpuKernelSources.java:808-841); a candidate
// it rejects because an earlier ref of the same leaf got closer cannot be the
// minimum, so the winner is the same, and the counters, which count tests
// rather than acceptances, are the same.

constexpr int kRegroupPerLane = 8;                       // refs per lane per round
constexpr int kTaskCap = 64 * kRegroupPerLane;

struct RegroupLds {
	float ray[11][64];      // o, d, 1/d, dot(d, d), 1/dot(d, d); then the fast flag
	uint8_t fast[64];
	unsigned long long key[64];
	float limitT[64];
	uint32_t taskRef[kTaskCap];
	uint8_t taskOwner[kTaskCap];
	uint32_t typeCount[4];
	uint32_t typeCursor[4];
};

__device__ __forceinline__ void waveSync() {
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	__builtin_amdgcn_wave_barrier();
}
// (t, rank) as one unsigned 64-bit key ordered like the pair.
__device__ __forceinline__ unsigned long long bestKey(float t, uint32_t rank) {
	uint32_t b = __float_as_uint(t);
	b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
	return ((unsigned long long) b << 32) | rank;
}
__device__ __forceinline__ float keyT(unsigned long long k) {
	uint32_t b = (uint32_t) (k >> 32);
	b = (b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b;
	return __uint_as_float(b);
}

template <bool kMetrics, int kCap, int kShort>
__global__ __launch_bounds__(kBlock) void wfTraceRegroup(WfParams P, int cur) {
	__shared__ StackLdsT<kShort> stackLds;
	__shared__ RegroupLds regroup[kWaves];
	const LdsStack lds = laneStack(stackLds);
	RegroupLds &G = regroup[threadIdx.x >> 6];
	const uint32_t count = P.counters[kCount0 + cur];
	const WfPool Q = P.pool[cur];
	const int lane = threadIdx.x & 63;
	Counters<kMetrics> c;
	c.zero();
	for (;;) {
		const uint32_t base = fetchChunk(P.counters + kWorkTrace);
		if (base >= count) break;
		const uint32_t i = base + lane;
		bool valid = i < count;
		uint4 meta = make_uint4(0, 0, 0, 0);
		if (valid) {
			meta = Q.meta[i];
			if ((meta.z >> 16) == kDead) {
				P.hitB[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				valid = false;
			}
		}
		Ray ray = makeRay(mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 1.0f), false);
		const SceneRange *Rp = P.scenes ? &P.scenes[valid ? (meta.z & 0xFFFFu) : 0].R : nullptr;
		Best best;
		best.t = 1e30f;
		best.rank = 0xFFFFFFFFu;
		best.kind = -1;
		best.index = -1;
		TravStack<kCap, kShort> st;
		st.lds = lds;
		int size = 0;
		bool overflow = false, trav = false;
		if (valid) {
			const float4 a = poolOrigin(P, Q, i, meta), b = Q.dir[i];
			ray = makeRay(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), Rp->fastSlabs != 0);
			c.add(0, 1);
			const SceneRange &R = *Rp;
			if (R.useBvh == 0) {
				traceLinear<kMetrics>(P.S, R, ray, best, c);
			} else if (R.refCount > 0 && R.rootCount > 0 && R.rootCount <= 4) {
				trav = true;
				for (int r = 0; r < R.rootCount; r++) {     // root insertion, as traceBvh
					const int ri = R.roots[r];
					if (ri < 0) { overflow = true; break; }
					c.add(1, 1);
					const float *rb = R.rootBounds[r];
					float e = entryDistance(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], ray, best.t);
					if (!isfinite(e)) continue;
					if (size >= R.stackLimit) { overflow = true; break; }
					int pos = size;
					while (pos > 0 && e >= st.entry(pos - 1)) {
						st.set(pos, st.node(pos - 1), st.entry(pos - 1));
						pos--;
					}
					st.set(pos, ri, e);
					size++;
					c.maxStack(size);
				}
			}
		}
		G.ray[0][lane] = ray.o.x; G.ray[1][lane] = ray.o.y; G.ray[2][lane] = ray.o.z;
		G.ray[3][lane] = ray.d.x; G.ray[4][lane] = ray.d.y; G.ray[5][lane] = ray.d.z;
		G.ray[6][lane] = ray.inv.x; G.ray[7][lane] = ray.inv.y; G.ray[8][lane] = ray.inv.z;
		G.ray[9][lane] = ray.dd; G.ray[10][lane] = ray.invDD;
		G.fast[lane] = ray.fast ? 1 : 0;
		for (;;) {
			// traverse interior nodes until this lane reaches a leaf or finishes
			bool atLeaf = false;
			int leafFirst = 0, leafCount = 0, leafKind = -1;
			while (trav && !overflow) {
				if (size == 0) { trav = false; break; }
				--size;
				int node;
				float entry;
				st.get(size, node, entry);
				const float ch = isfinite(best.t) ? fmaf(best.t, 1.0000015f, 1.0e-6f) : best.t;
				if (entry > ch) continue;
				const float4 *rec = P.S.nodes + 4 * (size_t) node;
				const int4 nd = *reinterpret_cast<const int4 *>(rec + 3);
				if (nd.x < 0) {
					c.add(11, 1);
					c.add(nd.x == -1 ? 13 : 12, 1);
					c.add(2, (uint32_t) nd.w);
					atLeaf = true;
					leafFirst = nd.z;
					leafCount = nd.w;
					leafKind = nd.x;
					break;
				}
				c.add(10, 1);
				const float4 a = rec[0], b = rec[1], cc = rec[2];
				const float le = entryDistance(a.x, a.y, a.z, a.w, b.x, b.y, ray, best.t);
				const float re = entryDistance(b.z, b.w, cc.x, cc.y, cc.z, cc.w, ray, best.t);
				c.add(1, 2);
				int nearN = nd.x, farN = nd.y;
				float nearE = le, farE = re;
				if (re < le) { nearN = nd.y; farN = nd.x; nearE = re; farE = le; }
				const int cap = Rp->stackLimit;
				if (isfinite(farE)) {
					if (size >= cap) { overflow = true; break; }
					st.set(size, farN, farE);
					size++;
					c.maxStack(size);
				}
				if (isfinite(nearE)) {
					if (size >= cap) { overflow = true; break; }
					st.set(size, nearN, nearE);
					size++;
					c.maxStack(size);
				}
			}
			if (__ballot(atLeaf) == 0ull) break;
			// leaf tests of the whole wave, regrouped by primitive class
			G.key[lane] = bestKey(best.t, best.rank);
			G.limitT[lane] = best.t;
			const bool mixed = leafKind > -2 || leafKind < -5;
			const uint32_t leafType = leafKind == -2 ? kRefSphere : leafKind == -3 ? kRefBox
			                        : leafKind == -4 ? kRefAffineSphere : kRefAffineBox;
			for (int round = 0;; round += kRegroupPerLane) {
				const int mine = atLeaf ? max(0, min(kRegroupPerLane, leafCount - round)) : 0;
				if (__ballot(mine > 0) == 0ull) break;
				if (lane < 4) G.typeCount[lane] = 0;
				waveSync();
				for (int k = 0; k < mine; k++) {
					const int ri = leafFirst + round + k;
					if (ri < Rp->refBase || ri >= Rp->refBase + Rp->refCount) continue;
					const uint32_t ref = P.S.refs[ri];
					const uint32_t type = mixed ? (ref & kRefTypeMask) : leafType;
					c.add(3 + typeSlot(type), 1);
					atomicAdd(&G.typeCount[type >> 30], 1u);
				}
				waveSync();
				if (lane < 4) {
					uint32_t off = 0;
					for (int t = 0; t < lane; t++) off += G.typeCount[t];
					G.typeCursor[lane] = off;
				}
				waveSync();
				for (int k = 0; k < mine; k++) {
					const int ri = leafFirst + round + k;
					if (ri < Rp->refBase || ri >= Rp->refBase + Rp->refCount) continue;
					const uint32_t ref = P.S.refs[ri];
					const uint32_t type = mixed ? (ref & kRefTypeMask) : leafType;
					const uint32_t slot = atomicAdd(&G.typeCursor[type >> 30], 1u);
					G.taskRef[slot] = type | (ref & kRefIndexMask);
					G.taskOwner[slot] = (uint8_t) lane;
				}
				waveSync();
				uint32_t off = 0;
				for (int t = 0; t < 4; t++) {
					const uint32_t n = G.typeCount[t];
					for (uint32_t b0 = 0; b0 < n; b0 += 64) {
						const uint32_t j = b0 + lane;
						if (j < n) {
							const uint32_t ref = G.taskRef[off + j];
							const int owner = G.taskOwner[off + j];
							Ray rr;
							rr.o = mk(G.ray[0][owner], G.ray[1][owner], G.ray[2][owner]);
							rr.d = mk(G.ray[3][owner], G.ray[4][owner], G.ray[5][owner]);
							rr.inv = mk(G.ray[6][owner], G.ray[7][owner], G.ray[8][owner]);
							rr.dd = G.ray[9][owner];
							rr.invDD = G.ray[10][owner];
							rr.fast = G.fast[owner] != 0;
							const float limit = nextafterf(G.limitT[owner], INFINITY);
							const int index = (int) (ref & kRefIndexMask);
							float tc = 0.0f;
							bool found;
							if (t == 0) found = testSphere(P.S.spheres[index], rr, limit, tc);
							else if (t == 1) found = testBox(P.S.boxes[2 * index], P.S.boxes[2 * index + 1], rr, limit, tc);
							else if (t == 2) found = testAffineSphere(P.S.affS + 3 * index, rr, limit, tc);
							else found = testAffineBox(P.S.affB + 3 * index, rr, limit, tc);
							if (found) atomicMin(&G.key[owner], bestKey(tc, ref));
						}
					}
					off += n;
				}
				waveSync();
			}
			if (atLeaf) {
				const unsigned long long k = G.key[lane];
				const uint32_t rank = (uint32_t) k;
				if (rank != best.rank || keyT(k) != best.t) {
					best.t = keyT(k);
					best.rank = rank;
					const uint32_t type = rank & kRefTypeMask;
					best.kind = type == kRefSphere ? kKindSphere : type == kRefBox ? kKindBox
					          : type == kRefAffineSphere ? kKindAffineSphere : kKindAffineBox;
					best.index = (int) (rank & kRefIndexMask);
				}
			}
		}
		if (valid) {
			const SceneRange &R = *Rp;
			if (overflow) {
				c.add(8, 1);
				for (int r = R.refBase; r < R.refBase + R.refCount; r++) {
					const uint32_t ref = P.S.refs[r];
					c.add(2, 1);
					c.add(3 + typeSlot(ref & kRefTypeMask), 1);
					testRef(P.S, ref, ref & kRefTypeMask, ray, best);
				}
			}
			c.add(2, (uint32_t) R.planeCount);
			c.add(5, (uint32_t) R.planeCount);
			tracePlanes(P.S, R, ray.o, ray.d, best);
			writeHit(P, i, ray.o, ray.d, best, best.kind >= 0);
		}
	}
	if (P.stats) flushCounters<kMetrics>(c, P.stats);
}

// Trace kernels compiled ahead of time: stack storage of 32, 64 or 256
// entries per lane (the autotuned stack depth; at least the runtime limit),
// of which 4 or 10 are held in LDS (the autotuned LDS split); optionally
// with primitive-class regrouping of the leaf tests.
using TraceFn = void (*)(WfParams, int);
// Storage entries for a configuration: the requested size, else 64 (256
// above a limit of 64).
inline int stackStorage(const mrad_wavefront_config &c) {
	return c.stack_storage > 0 ? c.stack_storage : (c.bvh_stack_size <= 64 ? 64 : 256);
}
// Shading fused into the trace kernel (wfTrace's kFuse) instead of the
// separate wfShade pass; the regrouped trace keeps the separate pass.
#ifndef MRAD_FUSED_SHADE
#define MRAD_FUSED_SHADE 0
#endif
constexpr bool kFusedShade = MRAD_FUSED_SHADE != 0;
inline bool fusedShade(const mrad_wavefront_config &c) { return kFusedShade && c.leaf_regroup == 0; }
template <bool M>
TraceFn traceFor(int storage, int ldsEntries, bool regroup) {
	constexpr bool F = kFusedShade;
	if (regroup) return storage <= 64 ? wfTraceRegroup<M, 64, 8> : wfTraceRegroup<M, 256, 8>;
	if (ldsEntries <= 4)
		return storage <= 32 ? wfTrace<M, 32, 4, F> : storage <= 64 ? wfTrace<M, 64, 4, F> : wfTrace<M, 256, 4, F>;
	return storage <= 32 ? wfTrace<M, 32, kShortStack, F>
	     : storage <= 64 ? wfTrace<M, 64, kShortStack, F> : wfTrace<M, 256, kShortStack, F>;
}

// The capped passes (trace_passes > 1) of the plain, unfused trace.
#ifndef MRAD_TRACE_PASSES
#define MRAD_TRACE_PASSES 2
#endif
#ifndef MRAD_PASS_POPS
#define MRAD_PASS_POPS 80
#endif
constexpr int kDefaultTracePasses = MRAD_TRACE_PASSES;
constexpr int kTwoPassPool = 1 << 25;     // smallest queue capacity given the default passes
constexpr int kDefaultPassPops = MRAD_PASS_POPS;
#ifndef MRAD_PASS_DENSE_LANES
#define MRAD_PASS_DENSE_LANES 16
#endif
constexpr int kDefaultPassDenseLanes = MRAD_PASS_DENSE_LANES;
using TracePassFn = void (*)(WfParams, int, int, int, int);
inline bool cappedPasses(const mrad_wavefront_config &c) {
	return c.trace_passes > 1 && c.leaf_regroup == 0 && !fusedShade(c);
}
template <bool M>
TracePassFn tracePassFor(int storage, int ldsEntries) {
	if (ldsEntries <= 4)
		return storage <= 32 ? wfTracePass<M, 32, 4> : storage <= 64 ? wfTracePass<M, 64, 4> : wfTracePass<M, 256, 4>;
	return storage <= 32 ? wfTracePass<M, 32, kShortStack>
	     : storage <= 64 ? wfTracePass<M, 64, kShortStack> : wfTracePass<M, 256, kShortStack>;
}

}  // namespace mrad

using namespace mrad;

struct mrad_batch {
	std::mutex mu;
	mrad_wavefront_config cfg{};
	bool available = false;
	std::string unavailableReason;
	hipStream_t stream = nullptr;
	hipEvent_t evStart = nullptr, evEnd = nullptr;
	int cuCount = 0;
	int shadeBlocksPerCu = 1;           // resident shade blocks per CU
	HostArena host;
	std::vector<WfScene> scenes;
	bool dirty = true;
	DevArena arena;
	DevBuf sceneTable, chunkTable, pool[2][4], hitA, hitB, results, images, counters, stats;
	DevBuf contHdr[2], contStack[2];     // continuation lists of the capped trace passes
	hipGraphExec_t graph = nullptr;
	std::vector<hipEvent_t> stepEvents;
	uint64_t totalPaths = 0, totalPixels = 0;
	int maxSpp = 1;
	std::vector<WfChunk> chunks;
	uint32_t chunkPaths = 0;             // result slots per buffer of the current plan
	uint32_t *hostCounters = nullptr;    // pinned: counter block read back after each graph launch
	mrad_batch_stats last{};
	// a batch of one view over an arena owned elsewhere (the tracer's resident
	// scene, wavefrontFrame): no arena upload of its own
	bool external = false;
	DevScene externalView{};
};

namespace {

int requireBatch(mrad_batch *b) {
	if (!b) return set_error(MRAD_ERR_INVALID_ARGUMENT, "batch handle is null");
	if (!b->available) return set_error(MRAD_ERR_UNAVAILABLE, "HIP batch renderer is unavailable: " + b->unavailableReason);
	return MRAD_OK;
}

void dropGraph(mrad_batch *b) {
	if (b->graph) (void) hipGraphExecDestroy(b->graph);
	b->graph = nullptr;
}

// Result slots per buffer for a configuration: the requested number, or 8
// pools' worth clamped to [2^24, 2^30]; never fewer than one pixel's samples.
uint32_t chunkPathsFor(const mrad_wavefront_config &c, int maxSpp) {
	uint64_t n = c.result_chunk_paths > 0 ? (uint64_t) c.result_chunk_paths
	                                      : std::min<uint64_t>(1ull << 30, std::max<uint64_t>(1ull << 24, 8ull * c.queue_capacity));
	return (uint32_t) std::max<uint64_t>(n, (uint64_t) maxSpp);
}

// Cuts the batch into chunks of whole pixels of at most chunkPaths paths each.
void planChunks(mrad_batch *b) {
	const uint32_t cp = b->chunkPaths;
	b->chunks.clear();
	WfChunk cur{0, 0, 0, 0};
	for (const WfScene &V : b->scenes) {
		const uint64_t npx = (uint64_t) V.width * (uint64_t) V.height;
		uint64_t px = 0;
		while (px < npx) {
			const uint64_t room = (cp - cur.pathCount) / (uint32_t) V.spp;
			if (room == 0) {
				b->chunks.push_back(cur);
				cur = WfChunk{cur.pathStart + cur.pathCount, cur.pixelStart + cur.pixelCount, 0, 0};
				continue;
			}
			const uint64_t take = std::min(room, npx - px);
			cur.pathCount += (uint32_t) (take * (uint32_t) V.spp);
			cur.pixelCount += (uint32_t) take;
			px += take;
		}
	}
	if (cur.pathCount > 0) b->chunks.push_back(cur);
}

// Uploads the arena, scene and chunk tables and sizes the pools and result
// buffers for the current scenes and configuration.
int prepare(mrad_batch *b) {
	if (!b->dirty) return MRAD_OK;
	dropGraph(b);
	if (!b->external) MRAD_TRY(b->arena.upload(b->stream, b->host));
	MRAD_TRY(upload(b->stream, b->sceneTable, b->scenes));
	b->chunkPaths = chunkPathsFor(b->cfg, b->maxSpp);
	planChunks(b);
	MRAD_TRY(upload(b->stream, b->chunkTable, b->chunks));
	const uint32_t cap = (uint32_t) b->cfg.queue_capacity;
	for (int q = 0; q < 2; q++)
		for (int k = 0; k < 4; k++) MRAD_TRY(ensure(b->pool[q][k], (size_t) cap * 16));
	MRAD_TRY(ensure(b->hitA, (size_t) cap * 16));
	MRAD_TRY(ensure(b->hitB, (size_t) cap * 16));
	// buffer 0 holds the even chunks, buffer 1 (from slot chunkPaths) the odd ones
	const size_t slots = b->chunks.size() > 1 ? 2 * (size_t) b->chunkPaths : (size_t) b->chunks[0].pathCount;
	MRAD_TRY(ensure(b->results, slots * 12));
	MRAD_TRY(ensure(b->images, (size_t) b->totalPixels * 12));
	MRAD_TRY(ensure(b->counters, kCounterWords * 4));
	MRAD_TRY(ensure(b->stats, MRAD_STAT_COUNT * 8));
	MRAD_HIP(hipStreamSynchronize(b->stream));
	b->dirty = false;
	return MRAD_OK;
}

// Continuation lists of the capped trace passes, sized for the configuration.
// A buffer that moves (or is released by a failed allocation) drops the
// captured graph, which holds the old address.
int ensureSuspend(mrad_batch *b) {
	if (!cappedPasses(b->cfg)) return MRAD_OK;
	const size_t n = (size_t) b->cfg.suspend_capacity;
	int rc = MRAD_OK;
	for (int q = 0; q < 2 && rc == MRAD_OK; q++) {
		const void *h = b->contHdr[q].p, *st = b->contStack[q].p;
		rc = ensure(b->contHdr[q], n * 16);
		if (rc == MRAD_OK) rc = ensure(b->contStack[q], n * kSuspendEntries * 8);
		if (b->contHdr[q].p != h || b->contStack[q].p != st) dropGraph(b);
	}
	return rc;
}

WfParams params(mrad_batch *b) {
	WfParams P{};
	P.S = b->external ? b->externalView : b->arena.view;
	P.scenes = b->sceneTable.as<const WfScene>();
	P.sceneCount = (int) b->scenes.size();
	for (int q = 0; q < 2; q++) {
		P.pool[q].ori = b->pool[q][0].as<float4>();
		P.pool[q].dir = b->pool[q][1].as<float4>();
		P.pool[q].tl = b->pool[q][2].as<float4>();
		P.pool[q].meta = b->pool[q][3].as<uint4>();
	}
	P.hitA = b->hitA.as<float4>();
	P.hitB = b->hitB.as<float4>();
	P.results = b->results.as<float>();
	P.images = b->images.as<float>();
	P.counters = b->counters.as<uint32_t>();
	P.chunks = b->chunkTable.as<const WfChunk>();
	P.chunkCount = (uint32_t) b->chunks.size();
	P.chunkPaths = b->chunkPaths;
	P.capacity = (uint32_t) b->cfg.queue_capacity;
	P.stats = b->cfg.collect_metrics ? b->stats.as<unsigned long long>() : nullptr;
	for (int q = 0; q < 2; q++) {
		P.cont[q].hdr = b->contHdr[q].as<uint4>();
		P.cont[q].stack = b->contStack[q].as<unsigned long long>();
	}
	P.contCap = cappedPasses(b->cfg) ? (uint32_t) b->cfg.suspend_capacity : 0u;
	return P;
}

// One bounce step on pool `cur`: generate, trace, shade into cur ^ 1 (the
// shading inside the trace kernel when it is fused).  With `ev` (4 events) the
// kernel boundaries are recorded for per-kernel timing.
int enqueueStep(mrad_batch *b, const WfParams &P, int cur, hipEvent_t *ev = nullptr) {
	const int persistentBlocks = std::max(1, b->cuCount * b->cfg.waves_per_cu / kWaves);
	const int traceBlocks = std::max(1, b->cuCount * b->cfg.waves_per_cu / kTraceWaves);
	// shading takes fixed tiles: one resident wave of blocks, whatever the
	// trace kernel's setting
	const int shadeBlocks = std::max(1, b->cuCount * b->shadeBlocksPerCu);
	const int genBlocks = std::max(1, std::min(8192, (int) ((b->cfg.queue_capacity + kBlock - 1) / kBlock)));
	if (ev) MRAD_HIP(hipEventRecord(ev[0], b->stream));
	hipLaunchKernelGGL(wfGenerate, dim3(genBlocks), dim3(kBlock), 0, b->stream, P, cur);
	hipLaunchKernelGGL(wfAdvanceGenerate, dim3(1), dim3(64), 0, b->stream, P, cur);
	if (ev) MRAD_HIP(hipEventRecord(ev[1], b->stream));
	const bool regroup = b->cfg.leaf_regroup != 0;
	const TraceFn trace = b->cfg.collect_metrics
	                          ? traceFor<true>(stackStorage(b->cfg), b->cfg.lds_stack_entries, regroup)
	                          : traceFor<false>(stackStorage(b->cfg), b->cfg.lds_stack_entries, regroup);
	if (cappedPasses(b->cfg)) {
		const TracePassFn pass = b->cfg.collect_metrics
		                             ? tracePassFor<true>(stackStorage(b->cfg), b->cfg.lds_stack_entries)
		                             : tracePassFor<false>(stackStorage(b->cfg), b->cfg.lds_stack_entries);
		for (int p = 0; p < b->cfg.trace_passes; p++) {
			const int budget = p + 1 < b->cfg.trace_passes ? b->cfg.pass_pops : 0;
			hipLaunchKernelGGL(pass, dim3(traceBlocks), dim3(kTraceBlock), 0, b->stream, P, cur, p, budget,
			                   b->cfg.pass_dense_lanes);
		}
	} else {
		if (regroup) hipLaunchKernelGGL(trace, dim3(persistentBlocks), dim3(kBlock), 0, b->stream, P, cur);
		else hipLaunchKernelGGL(trace, dim3(traceBlocks), dim3(kTraceBlock), 0, b->stream, P, cur);
	}
	if (ev) MRAD_HIP(hipEventRecord(ev[2], b->stream));
	if (!fusedShade(b->cfg)) hipLaunchKernelGGL(wfShade, dim3(shadeBlocks), dim3(kBlock), 0, b->stream, P, cur);
	hipLaunchKernelGGL(wfAdvanceShade, dim3(1), dim3(64), 0, b->stream, P, cur);
	if (ev) MRAD_HIP(hipEventRecord(ev[3], b->stream));
	MRAD_HIP(hipGetLastError());
	return MRAD_OK;
}

// Adds the kernel intervals of one timed step to the stats.
int accumulateStep(mrad_batch *b, hipEvent_t *ev) {
	float g = 0.0f, tr = 0.0f, sh = 0.0f;
	MRAD_HIP(hipEventSynchronize(ev[3]));
	MRAD_HIP(hipEventElapsedTime(&g, ev[0], ev[1]));
	MRAD_HIP(hipEventElapsedTime(&tr, ev[1], ev[2]));
	MRAD_HIP(hipEventElapsedTime(&sh, ev[2], ev[3]));
	b->last.generate_nanos += (int64_t) (g * 1.0e6f);
	b->last.trace_nanos += (int64_t) (tr * 1.0e6f);
	b->last.shade_nanos += (int64_t) (sh * 1.0e6f);
	return MRAD_OK;
}

// Defaults and limits of a wavefront configuration (create and configure).
int normalizeConfig(mrad_wavefront_config &c) {
	if (c.queue_capacity == 0) c.queue_capacity = 1 << 22;
	if (c.waves_per_cu == 0) c.waves_per_cu = 28;
	if (c.bvh_stack_size == 0) c.bvh_stack_size = 32;
	if (c.lds_stack_entries == 0) c.lds_stack_entries = kShortStack;
	if (c.lds_stack_entries != 4 && c.lds_stack_entries != kShortStack)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "LDS stack entries must be 4 or " + std::to_string(kShortStack));
	if (c.queue_capacity < 64 || c.queue_capacity > (1 << 28))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "queue capacity must be in [64, 2^28]");
	if (c.waves_per_cu < 1 || c.waves_per_cu > 32) return set_error(MRAD_ERR_INVALID_ARGUMENT, "waves per CU must be in [1, 32]");
	if (c.bvh_stack_size < 8 || c.bvh_stack_size > 256)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "BVH stack size must be in [8, 256]");
	if (c.stack_storage != 0 && c.stack_storage != 32 && c.stack_storage != 64 && c.stack_storage != 256)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "stack storage must be 0, 32, 64 or 256 entries");
	if (c.stack_storage != 0 && c.stack_storage < c.bvh_stack_size)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "stack storage must hold the BVH stack size");
	if (c.result_chunk_paths < 0 || c.result_chunk_paths > (1 << 30))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "result chunk paths must be in [0, 2^30]");
	c.queue_capacity = (c.queue_capacity + 63) & ~63;
	// capped passes pay where a bounce step holds many paths; with a small pool
	// the steps are many and short, and the second launch and its tail cost more
	// than the regrouping saves (one pass: het-100k at 2^22 paths 12 % faster)
	if (c.trace_passes == 0) c.trace_passes = c.queue_capacity >= kTwoPassPool ? kDefaultTracePasses : 1;
	if (c.pass_pops == 0) c.pass_pops = kDefaultPassPops;
	if (c.suspend_capacity == 0) c.suspend_capacity = std::max(64, c.queue_capacity / 2);
	if (c.pass_dense_lanes == 0) c.pass_dense_lanes = kDefaultPassDenseLanes;
	if (c.trace_passes < 1 || c.trace_passes > 3)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "trace passes must be 1, 2 or 3");
	if (c.pass_pops < 1 || c.pass_pops > (1 << 20))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "pass pops must be in [1, 2^20]");
	if (c.suspend_capacity < 64 || c.suspend_capacity > c.queue_capacity)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "suspend capacity must be in [64, queue capacity]");
	if (c.pass_dense_lanes != -1 && (c.pass_dense_lanes < 1 || c.pass_dense_lanes > 64))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "pass dense lanes must be -1 or in [1, 64]");
	return MRAD_OK;
}

int buildGraph(mrad_batch *b, const WfParams &P) {
	if (b->graph) return MRAD_OK;
	hipGraph_t g = nullptr;
	MRAD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
	int rc = enqueueStep(b, P, 0);
	if (rc == MRAD_OK) rc = enqueueStep(b, P, 1);
	hipError_t e = hipStreamEndCapture(b->stream, &g);
	if (rc != MRAD_OK) {
		if (g) (void) hipGraphDestroy(g);
		return rc;
	}
	if (e != hipSuccess) return set_error(MRAD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
	e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
	(void) hipGraphDestroy(g);
	if (e != hipSuccess) {
		b->graph = nullptr;
		return set_error(MRAD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
	}
	return MRAD_OK;
}

// The render of mrad_batch_render; the caller holds the batch's lock with its
// device current and the batch holds at least one scene.
int renderLocked(mrad_batch *b) {
	const int64_t t0 = nowNanos();
	const bool wasDirty = b->dirty;
	MRAD_TRY(prepare(b));
	MRAD_TRY(ensureSuspend(b));
	b->last.upload_nanos = wasDirty ? nowNanos() - t0 : 0;
	const WfParams P = params(b);
	const bool timed = b->cfg.time_kernels != 0;
	if (b->cfg.use_graph && !timed) MRAD_TRY(buildGraph(b, P));
	b->last.trace_nanos = b->last.shade_nanos = b->last.generate_nanos = 0;
	while (timed && b->stepEvents.size() < 4) {
		hipEvent_t e;
		MRAD_HIP(hipEventCreate(&e));
		b->stepEvents.push_back(e);
	}
	// counters: zero, chunks 0 and 1 released (a kernel, not a memset or copy)
	const uint32_t chunkCount = (uint32_t) b->chunks.size();
	hipLaunchKernelGGL(wfInitCounters, dim3(1), dim3(64), 0, b->stream, P);
	MRAD_HIP(hipGetLastError());
	if (b->cfg.collect_metrics) MRAD_HIP(hipMemsetAsync(b->stats.p, 0, MRAD_STAT_COUNT * 8, b->stream));
	MRAD_HIP(hipEventRecord(b->evStart, b->stream));
	uint32_t resolved = 0;        // chunks resolved so far, in order
	int64_t steps = 0;
	const int64_t stepGuard = 64 + 4 * 33 * ((int64_t) (b->totalPaths / 64) + chunkCount + 1);
	for (;;) {
		if (timed) {
			for (int cur = 0; cur < 2; cur++) {
				MRAD_TRY(enqueueStep(b, P, cur, b->stepEvents.data()));
				MRAD_TRY(accumulateStep(b, b->stepEvents.data()));
			}
		} else if (b->graph) {
			MRAD_HIP(hipGraphLaunch(b->graph, b->stream));
		} else {
			MRAD_TRY(enqueueStep(b, P, 0));
			MRAD_TRY(enqueueStep(b, P, 1));
		}
		steps += 2;
		MRAD_HIP(hipMemcpyAsync(b->hostCounters, b->counters.p, kCounterWords * 4, hipMemcpyDeviceToHost, b->stream));
		MRAD_HIP(hipStreamSynchronize(b->stream));
		// Every chunk whose paths have all finished is resolved, in order; its
		// buffer's count is reset and the chunk two ahead may then be generated
		// into it.  Both kernels are stream-ordered before the next graph launch.
		// Only chunks released before this snapshot can be complete: the count
		// word of chunk r + 2 still holds chunk r's count until the release.
		const uint32_t *c = b->hostCounters;
		const uint32_t before = resolved;
		while (resolved < chunkCount && resolved < c[kGenLimit] &&
		       c[kDone0 + (resolved & 1u)] == b->chunks[resolved].pathCount) {
			const uint32_t px = b->chunks[resolved].pixelCount;
			const int blocks = std::max(1, std::min(16384, (int) ((px + kBlock - 1) / kBlock)));
			hipLaunchKernelGGL(wfResolve, dim3(blocks), dim3(kBlock), 0, b->stream, P, resolved);
			MRAD_HIP(hipGetLastError());
			resolved++;
		}
		if (resolved == chunkCount) break;
		if (resolved != before) {
			hipLaunchKernelGGL(wfRelease, dim3(1), dim3(64), 0, b->stream, P, before, resolved);
			MRAD_HIP(hipGetLastError());
		}
		if (steps > stepGuard) {
			const uint32_t r = resolved;
			return set_error(MRAD_ERR_ILLEGAL_STATE,
			                 "wavefront did not drain: " + std::to_string(r) + " of " + std::to_string(chunkCount) +
			                     " chunks resolved (chunk " + std::to_string(r) + ": " +
			                     std::to_string(b->chunks[r].pathCount) + " paths, " +
			                     std::to_string(c[kDone0 + (r & 1u)]) + " finished); pools " + std::to_string(c[kCount0]) +
			                     " / " + std::to_string(c[kCount1]) + ", generating chunk " + std::to_string(c[kGenChunk]) +
			                     " at " + std::to_string(c[kGenOffset]) + " (limit " + std::to_string(c[kGenLimit]) +
			                     "), finished " + std::to_string(c[kDone0]) + " / " + std::to_string(c[kDone1]));
		}
	}
	MRAD_HIP(hipEventRecord(b->evEnd, b->stream));
	MRAD_HIP(hipEventSynchronize(b->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, b->evStart, b->evEnd));
	b->last.kernel_nanos = (int64_t) (ms * 1.0e6f);
	b->last.steps = steps;
	b->last.paths = b->totalPaths;
	if (b->cfg.collect_metrics) MRAD_HIP(hipMemcpy(b->last.counters, b->stats.p, sizeof b->last.counters, hipMemcpyDeviceToHost));
	else std::memset(b->last.counters, 0, sizeof b->last.counters);
	b->last.total_nanos = nowNanos() - t0;
	return MRAD_OK;
}

}  // namespace

extern "C" {

int mrad_batch_create(const mrad_wavefront_config *config, mrad_batch **out) {
	clear_error();
	if (!config || !out) return set_error(MRAD_ERR_INVALID_ARGUMENT, "config and output handle are required");
	*out = nullptr;
	mrad_wavefront_config c = *config;
	MRAD_TRY(normalizeConfig(c));
	mrad_batch *b = new (std::nothrow) mrad_batch();
	if (!b) return set_error(MRAD_ERR_OUT_OF_MEMORY, "batch allocation failed");
	b->cfg = c;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count <= 0) {
		b->unavailableReason = e != hipSuccess ? hipGetErrorString(e) : "no HIP device";
	} else if (c.device < 0 || c.device >= count) {
		b->unavailableReason = "device ordinal out of range";
	} else if ((e = hipSetDevice(c.device)) != hipSuccess ||
	           (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
	           (e = hipEventCreate(&b->evStart)) != hipSuccess || (e = hipEventCreate(&b->evEnd)) != hipSuccess ||
	           (e = hipDeviceGetAttribute(&b->cuCount, hipDeviceAttributeMultiprocessorCount, c.device)) != hipSuccess ||
	           (e = hipHostMalloc((void **) &b->hostCounters, kCounterWords * 4, hipHostMallocDefault)) != hipSuccess) {
		b->unavailableReason = hipGetErrorString(e);
	} else if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->shadeBlocksPerCu, wfShade, kBlock, 0)) !=
	           hipSuccess) {
		b->unavailableReason = hipGetErrorString(e);
	} else {
		b->shadeBlocksPerCu = std::max(1, b->shadeBlocksPerCu);
		b->available = true;
	}
	*out = b;
	return MRAD_OK;
}

void mrad_batch_destroy(mrad_batch *b) {
	if (!b) return;
	{
		std::lock_guard<std::mutex> g(b->mu);
		if (b->available) {
			(void) hipSetDevice(b->cfg.device);
			if (b->stream) (void) hipStreamSynchronize(b->stream);
			dropGraph(b);
			b->arena.release();
			for (int q = 0; q < 2; q++)
				for (int k = 0; k < 4; k++) b->pool[q][k].release();
			for (DevBuf *x : {&b->sceneTable, &b->chunkTable, &b->hitA, &b->hitB, &b->results, &b->images, &b->counters,
			                  &b->stats})
				x->release();
			if (b->hostCounters) (void) hipHostFree(b->hostCounters);
			for (hipEvent_t e : b->stepEvents) (void) hipEventDestroy(e);
			if (b->evStart) (void) hipEventDestroy(b->evStart);
			if (b->evEnd) (void) hipEventDestroy(b->evEnd);
			if (b->stream) (void) hipStreamDestroy(b->stream);
		}
	}
	delete b;
}

int mrad_batch_add_scene(mrad_batch *b, const mrad_scene_arrays *a, const mrad_view *v, int *index) {
	clear_error();
	if (!b || !a || !v) return set_error(MRAD_ERR_INVALID_ARGUMENT, "batch, scene and view are required");
	if (v->width < 1 || v->height < 1 || v->samples < 1)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "view width, height and samples must be positive");
	MRAD_TRY(mrad_validate_full_path_depth(v->max_depth));
	if (v->rr_start < -1) return set_error(MRAD_ERR_INVALID_ARGUMENT, "roulette start depth must be >= -1");
	// a scene's pixel index is 32-bit (the reference's y * width + x); paths of
	// the whole batch are numbered with 64 bits
	const uint64_t pixels = (uint64_t) v->width * v->height;
	if (pixels > 0xFFFFFFFFull) return set_error(MRAD_ERR_INVALID_ARGUMENT, "view exceeds 2^32 pixels");
	if (v->samples > (1 << 30)) return set_error(MRAD_ERR_INVALID_ARGUMENT, "view samples exceed 2^30");
	const uint64_t paths = pixels * (uint64_t) v->samples;
	std::lock_guard<std::mutex> g(b->mu);
	if (b->scenes.size() >= 0xFFFF) return set_error(MRAD_ERR_INVALID_ARGUMENT, "at most 65535 scenes per batch");
	WfScene s{};
	MRAD_TRY(b->host.append(a, b->cfg.linear_scan == 0, b->cfg.bvh_stack_size, s.R));
	std::memcpy(s.cam, v->camera, sizeof s.cam);
	s.width = v->width;
	s.height = v->height;
	s.spp = v->samples;
	s.maxDepth = v->max_depth;
	s.rrStart = v->rr_start;
	s.frameIndex = v->frame_index;
	s.frameSeed = v->frame_seed;
	s.pathBase = b->totalPaths;
	s.pixelBase = b->totalPixels;
	b->totalPaths += paths;
	b->totalPixels += pixels;
	b->maxSpp = std::max(b->maxSpp, v->samples);
	b->scenes.push_back(s);
	b->dirty = true;
	if (index) *index = (int) b->scenes.size() - 1;
	return MRAD_OK;
}

int mrad_batch_render(mrad_batch *b) {
	clear_error();
	MRAD_TRY(requireBatch(b));
	std::lock_guard<std::mutex> g(b->mu);
	if (b->scenes.empty()) return set_error(MRAD_ERR_ILLEGAL_STATE, "the batch holds no scene");
	MRAD_HIP(hipSetDevice(b->cfg.device));
	return renderLocked(b);
}


// Changes the launch parameters of a batch (queue capacity, waves per CU, LDS
// split, graph use, timing, metrics) keeping its scenes resident; device,
// stack size and traversal mode must stay.  Used by the autotuner.
int mrad_batch_configure(mrad_batch *b, const mrad_wavefront_config *config) {
	clear_error();
	MRAD_TRY(requireBatch(b));
	if (!config) return set_error(MRAD_ERR_INVALID_ARGUMENT, "config is required");
	mrad_wavefront_config c = *config;
	MRAD_TRY(normalizeConfig(c));
	std::lock_guard<std::mutex> g(b->mu);
	if (c.device != b->cfg.device || c.bvh_stack_size != b->cfg.bvh_stack_size || c.linear_scan != b->cfg.linear_scan)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "device, stack size and traversal mode are fixed per batch");
	MRAD_HIP(hipSetDevice(b->cfg.device));
	// a new chunk size re-plans the chunks and result buffers on the next render
	if (chunkPathsFor(c, b->maxSpp) != chunkPathsFor(b->cfg, b->maxSpp)) b->dirty = true;
	dropGraph(b);
	const bool grow = (size_t) c.queue_capacity * 16 > b->pool[0][0].cap;
	if (grow && !b->dirty) {
		// The pools grow before the configuration changes.  ensure() frees a
		// buffer before allocating its replacement, so a failed grow leaves
		// some buffers released: the old configuration stays and the batch is
		// marked for a full prepare(), which reallocates them at its size.
		int rc = MRAD_OK;
		for (int q = 0; q < 2 && rc == MRAD_OK; q++)
			for (int k = 0; k < 4 && rc == MRAD_OK; k++) rc = ensure(b->pool[q][k], (size_t) c.queue_capacity * 16);
		if (rc == MRAD_OK) rc = ensure(b->hitA, (size_t) c.queue_capacity * 16);
		if (rc == MRAD_OK) rc = ensure(b->hitB, (size_t) c.queue_capacity * 16);
		if (rc != MRAD_OK) {
			b->dirty = true;
			return rc;
		}
	}
	b->cfg = c;
	return MRAD_OK;
}

int mrad_batch_copy_image(mrad_batch *b, int scene, float *rgb) {
	clear_error();
	MRAD_TRY(requireBatch(b));
	if (!rgb) return set_error(MRAD_ERR_INVALID_ARGUMENT, "RGB output must not be null");
	std::lock_guard<std::mutex> g(b->mu);
	if (scene < 0 || scene >= (int) b->scenes.size()) return set_error(MRAD_ERR_INVALID_ARGUMENT, "scene index out of range");
	if (b->dirty || !b->images.p) return set_error(MRAD_ERR_ILLEGAL_STATE, "the batch has not been rendered");
	const WfScene &s = b->scenes[scene];
	MRAD_HIP(hipSetDevice(b->cfg.device));
	MRAD_HIP(hipMemcpy(rgb, b->images.as<float>() + 3 * (size_t) s.pixelBase, (size_t) s.width * s.height * 12,
	                   hipMemcpyDeviceToHost));
	return MRAD_OK;
}

int mrad_batch_get_stats(mrad_batch *b, mrad_batch_stats *out) {
	clear_error();
	if (!b || !out) return set_error(MRAD_ERR_INVALID_ARGUMENT, "batch and output are required");
	std::lock_guard<std::mutex> g(b->mu);
	*out = b->last;
	return MRAD_OK;
}

int mrad_batch_get_config(mrad_batch *b, mrad_wavefront_config *out) {
	clear_error();
	if (!b || !out) return set_error(MRAD_ERR_INVALID_ARGUMENT, "batch and output are required");
	std::lock_guard<std::mutex> g(b->mu);
	*out = b->cfg;
	return MRAD_OK;
}

}  // extern "C"

namespace mrad {

// Pool size of the wavefront form of a frame: the frame's paths rounded up to
// whole waves, at most 2^23 (1.3 GB of pools and hit records).
static uint32_t framePool(uint64_t paths) {
	const uint64_t cap = std::min<uint64_t>(1ull << 23, std::max<uint64_t>(64, (paths + 63) & ~63ull));
	return (uint32_t) cap;
}

int wavefrontFrame(mrad_batch **batch, const mrad_tracer_config &cfg, const DevScene &S, const SceneRange &R,
                   const WavefrontView &v, float *out, mrad_batch_stats *stats) {
	const uint64_t pixels = (uint64_t) v.width * (uint64_t) v.height;
	const uint64_t paths = pixels * (uint64_t) v.spp;
	if (!*batch) {
		mrad_wavefront_config c{};
		c.device = cfg.device;
		c.queue_capacity = framePool(paths);
		c.bvh_stack_size = cfg.bvh_stack_size;
		c.collect_metrics = cfg.collect_metrics;
		c.use_graph = 1;
		MRAD_TRY(mrad_batch_create(&c, batch));
	}
	mrad_batch *b = *batch;
	MRAD_TRY(requireBatch(b));
	if (b->cfg.collect_metrics != cfg.collect_metrics) {
		mrad_wavefront_config c = b->cfg;
		c.collect_metrics = cfg.collect_metrics;
		MRAD_TRY(mrad_batch_configure(b, &c));
	}
	std::lock_guard<std::mutex> g(b->mu);
	MRAD_HIP(hipSetDevice(b->cfg.device));
	WfScene s{};
	s.R = R;
	std::memcpy(s.cam, v.cam, sizeof s.cam);
	s.width = v.width;
	s.height = v.height;
	s.spp = v.spp;
	s.maxDepth = v.maxDepth;
	s.rrStart = v.rrStart;
	s.frameIndex = v.frameIndex;
	s.frameSeed = v.frameSeed;
	const bool sameShape = b->scenes.size() == 1 && b->scenes[0].width == v.width &&
	                       b->scenes[0].height == v.height && b->scenes[0].spp == v.spp;
	const bool sameScene = b->external && std::memcmp(&b->externalView, &S, sizeof S) == 0 &&
	                       std::memcmp(&b->scenes[0].R, &R, sizeof R) == 0;
	b->external = true;
	b->externalView = S;
	if (!sameShape) {
		b->scenes.assign(1, s);
		b->totalPaths = paths;
		b->totalPixels = pixels;
		b->maxSpp = v.spp;
		b->dirty = true;
	} else {
		b->scenes[0] = s;
		// the graph reads the scene table from memory: a new camera, seed or
		// frame index needs no new graph, a new arena does (its pointers are
		// the graph's kernel arguments)
		if (!sameScene) dropGraph(b);
		if (!b->dirty) MRAD_TRY(upload(b->stream, b->sceneTable, b->scenes));
	}
	MRAD_TRY(renderLocked(b));
	MRAD_HIP(hipMemcpyAsync(out, b->images.p, (size_t) pixels * 12, hipMemcpyDeviceToDevice, b->stream));
	MRAD_HIP(hipStreamSynchronize(b->stream));
	*stats = b->last;
	return MRAD_OK;
}

}  // namespace mrad
