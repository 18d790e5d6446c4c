// Kernels and host runtime of the tracer entry points of include/mrad.h: the
// MI355X replacement of the reference's JCuda/NVRTC launcher
// (src/xyz/marsavic/gfxlab/gpu/GpuRayTracer.java) and of its two kernels,
// renderKernel and traceReplayKernel (gpu/GpuKernelSources.java:1167-1393).
// Code objects are compiled ahead of time for gfx950; there is no runtime
// compilation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mrad_host.hpp"
#include "mrad_kernels_common.hpp"
#include "mrad_tracer.hpp"

#ifndef MRAD_CODE_OBJECT_SHA256
#define MRAD_CODE_OBJECT_SHA256 "unavailable"
#endif

namespace mrad {

namespace {
thread_local std::string g_lastError;
}
int set_error(int code, const std::string &message) {
	g_lastError = message;
	return code;
}
void clear_error() { g_lastError.clear(); }
const char *last_error() { return g_lastError.c_str(); }

// ---------------------------------------------------------------- kernels

// Megakernel block: MRAD_RENDER_BLOCK threads in 8x8-pixel waves, 256 (a
// 16x16 tile of 4 waves) or 64 (one 8x8 wave per block).  A block's LDS is
// held until all its waves end, so with 4-wave blocks a CU waits for a tile's
// slowest wave before it starts another; one-wave blocks are refilled a wave
// at a time (4.5-6.6 % faster frames, same images).
#ifndef MRAD_RENDER_BLOCK
#define MRAD_RENDER_BLOCK 64
#endif
constexpr int kRenderBlock = MRAD_RENDER_BLOCK;
static_assert(kRenderBlock == 64 || kRenderBlock == 256, "render block is one or four waves");
constexpr int kTile = kRenderBlock == 256 ? 16 : 8;
// GpuLaunchProvenance.MAX_RENDER_PIXELS_PER_LAUNCH and GpuRayTracer's
// MAX_RENDER_LAUNCHES_PER_CALL (GpuLaunchProvenance.java:5, GpuRayTracer.java:550-564).
constexpr int kMaxPixelsPerLaunch = 262144;
constexpr long long kMaxRenderLaunches = 8192;
// The replay's launch plan (GpuRayTracer.java:64-76, :735-778): at most
// 262144 rays per launch; a linear replay also bounds each launch's
// primitive work (8 units per sphere, box or plane test, 64 per affine test)
// to 2^20 units per ray and the call to 8192 launches.
constexpr int kMaxReplayRaysPerLaunch = 262144;
constexpr long long kMaxLinearReplayWork = 1LL << 20;
constexpr long long kMaxLinearReplayChunks = 8192;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i mod 8).  Remapping i so that each XCD receives one contiguous range of
// tiles gives each XCD's private L2 a spatially coherent share of the image,
// but it also gives each XCD a band of very uneven cost (sky versus dense
// geometry); measured on the 100k field at 1080p this load imbalance costs
// 11 %, so the default keeps the round-robin spread.
#ifndef MRAD_XCD_REMAP
#define MRAD_XCD_REMAP 0
#endif
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdContiguous(int i, int n) {
	if (!MRAD_XCD_REMAP) return i;
	const int q = n / kXcds, r = n % kXcds;
	const int x = i % kXcds, k = i / kXcds;
	return x * q + min(x, r) + k;
}

struct RenderParams {
	DevScene S;
	SceneRange R;
	float cam[12];
	int width, height, maxDepth, rrStart, frameIndex, s0, s1, total, tilesX;
	int x0, y0, regionW, regionH;   // the pixel tile of this launch
	int laneShift;                  // log2 of the lanes that share a pixel (one-wave blocks)
	uint64_t frameSeed;
	float *out;
	unsigned long long *stats;
};

// Occupancy the megakernel is compiled for (waves per SIMD).
#ifndef MRAD_RENDER_WAVES_PER_EU
#define MRAD_RENDER_WAVES_PER_EU 6
#endif
// Node slots of the scene's top BVH levels each megakernel block stages in
// LDS at its start (0: none).
#ifndef MRAD_RENDER_TOP
#define MRAD_RENDER_TOP 0
#endif
constexpr int kRenderTop = MRAD_RENDER_TOP;
// A launch with fewer 8x8 tiles than MRAD_RENDER_TARGET_WAVES / k gives each
// pixel k lanes (k a power of two, at most 64 and at most the launch's
// samples): the k lanes trace consecutive samples side by side and the pixel's
// first lane adds their colours to its running sum in sample order, so the sums
// are those of one lane per pixel while a small frame fills k times the waves.
// 65,536 waves (1 = one lane per pixel always): 64 batch frames 1834 -> 1336 ms,
// het-100k 1080p 16 spp 42.2 -> 36.4 ms, the 512x512 32-spp forward 6.6 -> 2.5 ms
// (16,384 / 32,768 waves: 1580 / 1424 ms on the batch frames).
#ifndef MRAD_RENDER_TARGET_WAVES
#define MRAD_RENDER_TARGET_WAVES 65536
#endif
constexpr int kRenderTargetWaves = MRAD_RENDER_TARGET_WAVES;

// renderKernel (GpuKernelSources.java:1167-1315).  One thread per pixel runs
// samples [s0, s1) of the frame and keeps the running sum in registers; the
// float sum order and the last-sample normalization are the reference's.
// A wave covers an 8x8 pixel block for primary-ray coherence.  kBatch > 0
// tests popped leaves kBatch lanes at a time (traceBvh's leaf batching).
template <bool kMetrics, int kCap, int kBatch = 0, bool kSplit = false>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(MRAD_RENDER_WAVES_PER_EU)))
void renderKernel(RenderParams P) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int shift = kSplit ? P.laneShift : 0;   // P.laneShift > 0 only with one-wave blocks
	const int tile = xcdContiguous(blockIdx.x >> shift, gridDim.x >> shift);
	const int q = ((blockIdx.x & ((1 << shift) - 1)) << (6 - shift)) + (lane >> shift);   // pixel of the 8x8 block
	const int tileX = tile % P.tilesX, tileY = tile / P.tilesX;
	const int x = P.x0 + tileX * kTile + (wave & 1) * 8 + (q & 7);     // wave is 0 in a one-wave block
	const int y = P.y0 + tileY * kTile + (wave >> 1) * 8 + (q >> 3);
	const bool active = x < P.x0 + P.regionW && y < P.y0 + P.regionH;
	__shared__ StackLdsT<kShortStack, kRenderBlock / 64> stackLds;
	const LdsStack lds = laneStack(stackLds);
	constexpr int kTop = kBatch > 0 ? 0 : kRenderTop;
	__shared__ float4 topLds[kTop > 0 ? 4 * kTop : 1];
	TopNodes top{};
	if constexpr (kTop > 0) {
		const int n = P.R.useBvh ? min(kTop, P.R.topSlots) : 0;
		for (int i = threadIdx.x; i < 4 * n; i += kRenderBlock) topLds[i] = P.S.nodes[4 * (size_t) P.R.nodeBase + i];
		__syncthreads();
		top = TopNodes{(const LdsFloat4 *) topLds, P.R.nodeBase, n};
	}
	Counters<kMetrics> c;
	c.zero();
	if constexpr (kSplit) {
		// k lanes per pixel: round r traces samples s0 + r k + j (j = lane's
		// place in its pixel's group); the group's first lane adds the k colours
		// in sample order.  The shuffles run outside any divergent branch.
		const int k = 1 << shift, j = lane & (k - 1), first = lane & ~(k - 1);
		const uint32_t pixel = active ? (uint32_t) y * (uint32_t) P.width + (uint32_t) x : 0u;
		const size_t idx = (size_t) pixel * 3;
		f3 acc = active && j == 0 && P.s0 > 0 ? mk(P.out[idx], P.out[idx + 1], P.out[idx + 2]) : mk(0.0f, 0.0f, 0.0f);
		const int rounds = (P.s1 - P.s0 + k - 1) >> shift;
		for (int r = 0; r < rounds; r++) {
			const int s = P.s0 + (r << shift) + j;
			f3 col = mk(0.0f, 0.0f, 0.0f);
			if (active && s < P.s1) {
				uint32_t rng = path_seed(P.frameSeed, P.frameIndex, pixel, s);
				f3 o, d;
				cameraRay(P.cam, P.width, P.height, x, y, rng, o, d);
				col = tracePath<kMetrics, kCap, kBatch, kTop>(P.S, P.R, o, d, P.maxDepth, P.rrStart, rng, c, lds, top);
			}
			for (int jj = 0; jj < k; jj++) {
				const f3 cj = mk(__shfl(col.x, first + jj, 64), __shfl(col.y, first + jj, 64),
				                 __shfl(col.z, first + jj, 64));
				const int sj = P.s0 + (r << shift) + jj;
				if (j == 0 && sj < P.s1) {
					acc = sj == 0 ? cj : acc + cj;
					if (sj == P.total - 1) acc = acc * (1.0f / (float) P.total);
				}
			}
		}
		if (active && j == 0) {
			P.out[idx] = acc.x;
			P.out[idx + 1] = acc.y;
			P.out[idx + 2] = acc.z;
		}
	} else if (active) {
		const uint32_t pixel = (uint32_t) y * (uint32_t) P.width + (uint32_t) x;
		const size_t idx = (size_t) pixel * 3;
		f3 acc = P.s0 > 0 ? mk(P.out[idx], P.out[idx + 1], P.out[idx + 2]) : mk(0.0f, 0.0f, 0.0f);
		for (int s = P.s0; s < P.s1; s++) {
			uint32_t rng = path_seed(P.frameSeed, P.frameIndex, pixel, s);
			f3 o, d;
			cameraRay(P.cam, P.width, P.height, x, y, rng, o, d);
			const f3 col = tracePath<kMetrics, kCap, kBatch, kTop>(P.S, P.R, o, d, P.maxDepth, P.rrStart, rng, c, lds,
			                                                      top);
			acc = s == 0 ? col : acc + col;
			if (s == P.total - 1) acc = acc * (1.0f / (float) P.total);
		}
		P.out[idx] = acc.x;
		P.out[idx + 1] = acc.y;
		P.out[idx + 2] = acc.z;
	}
	if (P.stats) flushCounters<kMetrics>(c, P.stats);
}

struct ReplayParams {
	DevScene S;
	SceneRange R;
	const float *rays;
	int n;
	int *flags;
	float *dist;
	int *order;
	int *mat;
	float *normals;
	uint32_t *perRay;
	unsigned long long *stats;
	int matBase;
};

// traceReplayKernel (GpuKernelSources.java:1317-1393).
template <bool kMetrics, int kCap>
__global__ __launch_bounds__(kBlock) void replayKernel(ReplayParams P) {
	const int r = blockIdx.x * kBlock + threadIdx.x;
	__shared__ StackLds stackLds;
	const LdsStack lds = laneStack(stackLds);
	Counters<kMetrics> c;
	c.zero();
	if (r < P.n) {
		const float *ray = P.rays + 6 * (size_t) r;
		const f3 o = mk(ray[0], ray[1], ray[2]), d = mk(ray[3], ray[4], ray[5]);
		Best best;
		const bool found = traceClosest<kMetrics, kCap>(P.S, P.R, o, d, best, c, lds);
		c.add(0, 1);
		HitDetail h;
		if (found) h = resolveHit(P.S, o, d, best.t, best.kind, best.index);
		P.flags[r] = found ? 1 : 0;
		P.dist[r] = found ? best.t : INFINITY;
		P.order[r] = found ? localOrder(P.R, best) : -1;
		P.mat[r] = found ? (h.matIndex < 0 ? h.matIndex : h.matIndex - P.matBase) : -1;
		P.normals[3 * (size_t) r] = found ? h.normal.x : 0.0f;
		P.normals[3 * (size_t) r + 1] = found ? h.normal.y : 0.0f;
		P.normals[3 * (size_t) r + 2] = found ? h.normal.z : 0.0f;
		if constexpr (kMetrics) {
			if (P.perRay)
				for (int i = 0; i < kStatCount; i++) P.perRay[(size_t) r * kStatCount + i] = c.v[i];
		}
	}
	if (P.stats) flushCounters<kMetrics>(c, P.stats);
}

using RenderFn = void (*)(RenderParams);
using ReplayFn = void (*)(ReplayParams);
// Stack storage capacities compiled ahead of time; a limit up to 64 uses the
// 64-entry storage (LDS top + private spill), larger limits the 256 one.
// Leaf batching is compiled for one width, kLeafBatchLanes (16: the fastest
// of 8 / 16 / 32 on the Layered Hero frames).
#ifndef MRAD_LEAF_BATCH_LANES
#define MRAD_LEAF_BATCH_LANES 16
#endif
constexpr int kLeafBatchLanes = MRAD_LEAF_BATCH_LANES;
template <bool M, bool S>
RenderFn renderFor(int cap, int leafBatch) {
	if (leafBatch > 0)
		return cap <= 64 ? renderKernel<M, 64, kLeafBatchLanes, S> : renderKernel<M, 256, kLeafBatchLanes, S>;
	if (cap <= 64) return renderKernel<M, 64, 0, S>;
	return renderKernel<M, 256, 0, S>;
}
template <bool M>
ReplayFn replayFor(int cap) {
	if (cap <= 64) return replayKernel<M, 64>;
	return replayKernel<M, 256>;
}

}  // namespace mrad

using namespace mrad;

namespace {

// Device staging of one replay call: rays in; flags, distances, orders,
// material indices, normals, per-ray counters and the frame counters out.
struct ReplayBuffers {
	float *rays, *dist, *normals;
	int *flags, *order, *mat;
	uint32_t *perRay;
	unsigned long long *stats;
};

int replayChecks(mrad_tracer *t, const float *rays, int n, int32_t *flags, float *dist, int32_t *order, int32_t *mat,
                 float *normals, uint32_t *) {
	if (!rays || n <= 0) return set_error(MRAD_ERR_INVALID_ARGUMENT, "Replay rays must contain origin.xyz and direction.xyz.");
	if (!flags || !dist || !order || !mat || !normals) return set_error(MRAD_ERR_INVALID_ARGUMENT, "replay outputs must not be null");
	return requireTracer(t);
}

// Checks that need the handle's lock held.
int replayState(mrad_tracer *t, uint32_t *perRay) {
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident: call mrad_upload_scene first");
	if (perRay && !t->cfg.collect_metrics)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "per-ray counters need a tracer created with collect_metrics");
	return MRAD_OK;
}

int ensureEvents(mrad_tracer *t, int boundaries);

// Rays per physical replay launch (launchTraceReplayLocked and
// linearReplayRaysPerLaunch, GpuRayTracer.java:624-629, :743-778): the
// configured limit, and for a linear replay the primitive-work budget per
// launch and the chunk limit, with the reference's messages.
int replayPlan(const mrad_tracer *t, int n, int &perLaunch) {
	perLaunch = t->cfg.replay_rays_per_launch;
	if (t->cfg.bvh_traversal) return MRAD_OK;
	const SceneRange &R = t->range;
	const long long work = 8LL * ((long long) R.sphCount + R.boxCount + R.planeCount) +
	                       64LL * ((long long) R.asCount + R.abCount);
	if (work > kMaxLinearReplayWork)
		return set_error(MRAD_ERR_INVALID_ARGUMENT,
		                 "Linear replay requires " + std::to_string(work) +
		                     " primitive-work units per ray, exceeding the watchdog-safe budget of " +
		                     std::to_string(kMaxLinearReplayWork) + ". Use BVH traversal or a smaller scene.");
	if (work > 0) perLaunch = (int) std::min<long long>(perLaunch, kMaxLinearReplayWork / work);
	const long long chunks = ((long long) n + perLaunch - 1) / perLaunch;
	if (chunks > kMaxLinearReplayChunks)
		return set_error(MRAD_ERR_INVALID_ARGUMENT,
		                 "Linear replay would require " + std::to_string(chunks) +
		                     " watchdog-safe chunks, exceeding the limit of " + std::to_string(kMaxLinearReplayChunks) +
		                     ". Reduce the ray count or use BVH traversal.");
	return MRAD_OK;
}

// A replay on staging B: rays copied in, launched in chunks of the replay
// plan, the launches timed with the handle's events, results copied out,
// frame statistics updated.  With launchEvents the boundary of every launch
// is timed as well (events are created on first use); without, the longest
// launch is reported as the whole interval.  The caller holds the handle's
// lock with its device current.
int replayWith(mrad_tracer *t, const ReplayBuffers &B, const float *rays, int n, int collect, int32_t *flags,
               float *dist, int32_t *order, int32_t *mat, float *normals, uint32_t *perRay, bool launchEvents) {
	const int64_t t0 = nowNanos();
	const size_t N = (size_t) n;
	const bool metrics = collect && t->cfg.collect_metrics;
	int perLaunch = 0;
	MRAD_TRY(replayPlan(t, n, perLaunch));
	const int launches = (int) (((long long) n + perLaunch - 1) / perLaunch);
	// a BVH replay has no launch limit: beyond 8192 launches no boundary
	// events are kept (one event per launch would grow without bound)
	launchEvents = launchEvents && launches <= kMaxLinearReplayChunks;
	if (launchEvents) MRAD_TRY(ensureEvents(t, launches - 1));
	MRAD_HIP(hipMemcpyAsync(B.rays, rays, N * 24, hipMemcpyHostToDevice, t->stream));
	MRAD_HIP(hipMemsetAsync(B.stats, 0, MRAD_STAT_COUNT * 8, t->stream));
	ReplayParams P{};
	P.S = t->arena.view;
	P.R = t->range;
	const bool instrumented = metrics || perRay;
	P.stats = instrumented ? B.stats : nullptr;
	P.matBase = 0;
	ReplayFn fn = instrumented ? replayFor<true>(t->cfg.bvh_stack_size) : replayFor<false>(t->cfg.bvh_stack_size);
	MRAD_HIP(hipEventRecord(t->evStart, t->stream));
	for (int l = 0; l < launches; l++) {
		const size_t off = (size_t) l * (size_t) perLaunch;
		const int m = (int) std::min<size_t>((size_t) perLaunch, N - off);
		P.rays = B.rays + 6 * off;
		P.n = m;
		P.flags = B.flags + off;
		P.dist = B.dist + off;
		P.order = B.order + off;
		P.mat = B.mat + off;
		P.normals = B.normals + 3 * off;
		P.perRay = perRay ? B.perRay + MRAD_STAT_COUNT * off : nullptr;
		hipLaunchKernelGGL(fn, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, t->stream, P);
		MRAD_HIP(hipGetLastError());
		if (launchEvents && l + 1 < launches) MRAD_HIP(hipEventRecord(t->boundary[l], t->stream));
	}
	MRAD_HIP(hipEventRecord(t->evEnd, t->stream));
	MRAD_HIP(hipEventSynchronize(t->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, t->evStart, t->evEnd));
	int64_t maxPhys = (int64_t) (ms * 1.0e6f);
	if (launchEvents && launches > 1) {
		maxPhys = 0;
		hipEvent_t prev = t->evStart;
		for (int k = 0; k < launches; k++) {
			hipEvent_t end = k + 1 == launches ? t->evEnd : t->boundary[k];
			float pm = 0.0f;
			MRAD_HIP(hipEventElapsedTime(&pm, prev, end));
			maxPhys = std::max<int64_t>(maxPhys, (int64_t) (pm * 1.0e6f));
			prev = end;
		}
	}
	const int64_t c0 = nowNanos();
	MRAD_HIP(hipMemcpy(flags, B.flags, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(dist, B.dist, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(order, B.order, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(mat, B.mat, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(normals, B.normals, N * 12, hipMemcpyDeviceToHost));
	if (perRay) MRAD_HIP(hipMemcpy(perRay, B.perRay, N * MRAD_STAT_COUNT * 4, hipMemcpyDeviceToHost));
	if (instrumented) MRAD_HIP(hipMemcpy(t->last.counters, B.stats, sizeof t->last.counters, hipMemcpyDeviceToHost));
	else std::memset(t->last.counters, 0, sizeof t->last.counters);
	t->last.copy_nanos = nowNanos() - c0;
	t->last.kernel_nanos = (int64_t) (ms * 1.0e6f);
	t->last.maximum_physical_kernel_nanos = std::min(maxPhys, t->last.kernel_nanos);
	t->last.physical_launches = launches;
	t->last.total_nanos = nowNanos() - t0;
	return MRAD_OK;
}

// Scratch of mrad_trace_replay_noalloc, one copy per GPU for the life of the
// process: the 14 frame counters, then for up to MRAD_NOALLOC_MAX_RAYS rays
// the ray staging (24 B), flags, distances, orders, material indices (4 B
// each), normals (12 B) and per-ray counters (56 B); 6.8 MiB in all.  Every
// call on a GPU shares that GPU's copy.
constexpr size_t kNoAllocRays = MRAD_NOALLOC_MAX_RAYS;
constexpr size_t scratchAlign(size_t v) { return (v + 255) & ~(size_t) 255; }
constexpr size_t kScratchStats = 0;
constexpr size_t kScratchRays = scratchAlign(kScratchStats + MRAD_STAT_COUNT * 8);
constexpr size_t kScratchFlags = scratchAlign(kScratchRays + kNoAllocRays * 24);
constexpr size_t kScratchDist = scratchAlign(kScratchFlags + kNoAllocRays * 4);
constexpr size_t kScratchOrder = scratchAlign(kScratchDist + kNoAllocRays * 4);
constexpr size_t kScratchMat = scratchAlign(kScratchOrder + kNoAllocRays * 4);
constexpr size_t kScratchNormals = scratchAlign(kScratchMat + kNoAllocRays * 4);
constexpr size_t kScratchPerRay = scratchAlign(kScratchNormals + kNoAllocRays * 12);
constexpr size_t kScratchBytes = scratchAlign(kScratchPerRay + kNoAllocRays * MRAD_STAT_COUNT * 4);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The device address of `scratch` on each GPU.  In host code the array's
// name is only a host-side stand-in, so kernels and copies use the address
// hipGetSymbolAddress returns, taken once per device on the first call there
// (with that device current).
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMu;
unsigned char *g_scratchPtr[kMaxScratchDevices] = {};

int scratchFor(int device, unsigned char **out) {
	if (device < 0 || device >= kMaxScratchDevices)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "no-allocation scratch supports devices 0.." +
		                                                std::to_string(kMaxScratchDevices - 1));
	std::lock_guard<std::mutex> g(g_scratchMu);
	if (!g_scratchPtr[device]) {
		void *p = nullptr;
		MRAD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p) return set_error(MRAD_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
		g_scratchPtr[device] = static_cast<unsigned char *>(p);
	}
	*out = g_scratchPtr[device];
	return MRAD_OK;
}

int ensureEvents(mrad_tracer *t, int boundaries) {
	while ((int) t->boundary.size() < boundaries) {
		hipEvent_t e;
		MRAD_HIP(hipEventCreate(&e));
		t->boundary.push_back(e);
	}
	return MRAD_OK;
}

}  // namespace

extern "C" {

const char *mrad_last_error(void) { return mrad::last_error(); }

int mrad_validate_full_path_depth(int maxDepth) {
	clear_error();
	if (maxDepth < 0 || maxDepth > 32)
		return set_error(MRAD_ERR_INVALID_ARGUMENT,
		                 "GPU full-path depth must be in [0, 32] but was " + std::to_string(maxDepth));
	return MRAD_OK;
}

int mrad_code_object_sha256(char *out65) {
	clear_error();
	if (!out65) return set_error(MRAD_ERR_INVALID_ARGUMENT, "output buffer is null");
	std::strncpy(out65, MRAD_CODE_OBJECT_SHA256, 64);
	out65[64] = '\0';
	return MRAD_OK;
}

int mrad_create(const mrad_tracer_config *config, mrad_tracer **out) {
	clear_error();
	if (!config || !out) return set_error(MRAD_ERR_INVALID_ARGUMENT, "config and output handle are required");
	*out = nullptr;
	mrad_tracer_config c = *config;
	// GpuRayTracer.java:264-279: dimensions and spp clamp to >= 1.
	c.width = std::max(1, c.width);
	c.height = std::max(1, c.height);
	if ((long long) c.width * c.height > 0x7FFFFFFFLL / 3)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "GPU image dimensions exceed 32-bit indexing");
	c.samples_per_frame = std::max(1, c.samples_per_frame);
	if (c.bvh_stack_size == 0) c.bvh_stack_size = 32;
	if (c.bvh_stack_size < 8 || c.bvh_stack_size > 256)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "BVH stack size must be in [8, 256]");
	if (c.samples_per_launch <= 0 || c.samples_per_launch > c.samples_per_frame) c.samples_per_launch = c.samples_per_frame;
	// GpuLaunchProvenance.renderTileLayout (:44-60): 1 <= pixels per launch <= 262144
	if (c.pixels_per_launch < 0 || c.pixels_per_launch > kMaxPixelsPerLaunch)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "pixels per launch must be in [0, 262144]");
	if (c.replay_rays_per_launch == 0) c.replay_rays_per_launch = kMaxReplayRaysPerLaunch;
	if (c.replay_rays_per_launch < 1 || c.replay_rays_per_launch > kMaxReplayRaysPerLaunch)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "replay rays per launch must be in [0, 262144]");
	if (c.max_physical_kernel_nanos < 0)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "maximum physical kernel nanos must be >= 0");
	mrad_tracer *t = new (std::nothrow) mrad_tracer();
	if (!t) return set_error(MRAD_ERR_OUT_OF_MEMORY, "tracer allocation failed");
	t->cfg = c;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count <= 0) {
		t->unavailableReason = e != hipSuccess ? hipGetErrorString(e) : "no HIP device";
	} else if (c.device < 0 || c.device >= count) {
		t->unavailableReason = "device ordinal out of range";
	} else if ((e = hipSetDevice(c.device)) != hipSuccess ||
	           (e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking)) != hipSuccess ||
	           (e = hipEventCreate(&t->evStart)) != hipSuccess || (e = hipEventCreate(&t->evEnd)) != hipSuccess) {
		t->unavailableReason = hipGetErrorString(e);
	} else {
		t->available = true;
	}
	*out = t;
	return MRAD_OK;
}

void mrad_destroy(mrad_tracer *t) {
	if (!t) return;
	{
		std::lock_guard<std::mutex> g(t->mu);
		if (t->available) {
			(void) hipSetDevice(t->cfg.device);
			if (t->stream) (void) hipStreamSynchronize(t->stream);
			t->arena.release();
			for (DevBuf *b : {&t->output, &t->stats, &t->rays, &t->flags, &t->dist, &t->order, &t->mat, &t->normals,
			                  &t->perRay, &t->gradImage, &t->gradAlbedo, &t->gradEmission})
				b->release();
			for (hipEvent_t e : t->boundary) (void) hipEventDestroy(e);
			if (t->evStart) (void) hipEventDestroy(t->evStart);
			if (t->evEnd) (void) hipEventDestroy(t->evEnd);
			if (t->stream) (void) hipStreamDestroy(t->stream);
		}
		mrad_batch_destroy(t->wf);
		t->wf = nullptr;
	}
	delete t;
}

int mrad_is_available(mrad_tracer *t) { return t && t->available ? 1 : 0; }

int mrad_set_russian_roulette(mrad_tracer *t, int startDepth) {
	clear_error();
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	if (startDepth < -1) return set_error(MRAD_ERR_INVALID_ARGUMENT, "roulette start depth must be >= -1");
	std::lock_guard<std::mutex> g(t->mu);
	t->rrStart = startDepth;
	return MRAD_OK;
}

int mrad_set_leaf_batch(mrad_tracer *t, int lanes) {
	clear_error();
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	if (lanes != 0 && lanes != kLeafBatchLanes)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "leaf batch must be 0 or " + std::to_string(kLeafBatchLanes));
	std::lock_guard<std::mutex> g(t->mu);
	t->leafBatch = lanes;
	return MRAD_OK;
}

int mrad_set_render_form(mrad_tracer *t, int form) {
	clear_error();
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	if (form != MRAD_RENDER_FORM_MEGAKERNEL && form != MRAD_RENDER_FORM_WAVEFRONT)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "render form must be 0 (megakernel) or 1 (wavefront)");
	std::lock_guard<std::mutex> g(t->mu);
	t->renderForm = form;
	return MRAD_OK;
}

// ensureSceneResident / uploadScene (GpuRayTracer.java:873-948).
int mrad_upload_scene(mrad_tracer *t, const mrad_scene_arrays *a, uint64_t key) {
	clear_error();
	if (!t || !a) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer and GPU scene must not be null");
	MRAD_TRY(requireTracer(t));
	std::lock_guard<std::mutex> g(t->mu);
	MRAD_HIP(hipSetDevice(t->cfg.device));
	if (t->hasScene && t->residentKey == key && t->residentRevision == a->revision) {
		t->last.upload_nanos = 0;
		return MRAD_OK;
	}
	const int64_t t0 = nowNanos();
	t->hasScene = false;
	HostArena h;
	MRAD_TRY(h.append(a, t->cfg.bvh_traversal != 0, t->cfg.bvh_stack_size, t->range));
	MRAD_TRY(t->arena.upload(t->stream, h));
	t->materialCount = a->material_count;
	t->hasScene = true;
	t->residentKey = key;
	t->residentRevision = a->revision;
	t->last.upload_nanos = std::max<int64_t>(1, nowNanos() - t0);
	return MRAD_OK;
}

// The launch plan of one frame (renderTilePlan / renderLaunchCount,
// GpuRayTracer.java:544-564, with GpuLaunchProvenance.renderTileLayout): tiles
// of min(W, n) x max(1, n / tile width) pixels, one launch per tile and sample
// range; more than kMaxRenderLaunches launches is an invalid argument.  It
// needs no device, so callers check it before any upload.
struct LaunchPlan {
	int tileW, tileH, tileCols, tileRows, sampleLaunches;
	long long launches;
};

static int launchPlan(const mrad_tracer *t, LaunchPlan &L) {
	const int W = t->cfg.width, H = t->cfg.height, spp = t->cfg.samples_per_frame, per = t->cfg.samples_per_launch;
	const int ppl = t->cfg.pixels_per_launch;
	L.tileW = ppl > 0 ? std::min(W, ppl) : W;
	L.tileH = ppl > 0 ? std::min(H, std::max(1, ppl / L.tileW)) : H;
	L.tileCols = 1 + (W - 1) / L.tileW;
	L.tileRows = 1 + (H - 1) / L.tileH;
	L.sampleLaunches = (spp + per - 1) / per;
	L.launches = (long long) L.tileCols * L.tileRows * L.sampleLaunches;
	if (L.launches > kMaxRenderLaunches)
		return set_error(MRAD_ERR_INVALID_ARGUMENT,
		                 "Render would require " + std::to_string(L.launches) +
		                     " watchdog-safe pixel/sample launches, exceeding the limit of " +
		                     std::to_string(kMaxRenderLaunches) +
		                     ". Increase the bounded pixel cap or reduce samples per frame.");
	return MRAD_OK;
}

// renderLaunchCount (GpuRayTracer.java:556-564).
int mrad_render_launch_count(const mrad_tracer *t, int *launches) {
	clear_error();
	if (!t || !launches) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle and output must not be null");
	LaunchPlan L;
	MRAD_TRY(launchPlan(t, L));
	*launches = (int) L.launches;
	return MRAD_OK;
}

// renderSample in the wavefront form (mrad_set_render_form); the caller holds
// the handle's lock with its device current and a scene resident.  The frame's
// physical launches are its bounce steps; the longest one is not timed alone,
// so the whole kernel interval is reported as the longest.
static int renderWavefront(mrad_tracer *t, float *rgb, const float camera[12], int maxDepth, int frameIndex,
                           uint64_t frameSeed) {
	const int64_t frameStart = nowNanos();
	const int W = t->cfg.width, H = t->cfg.height;
	MRAD_TRY(ensure(t->output, (size_t) W * H * 3 * sizeof(float)));
	MRAD_HIP(hipStreamSynchronize(t->stream));          // the resident arena, on the handle's stream
	const WavefrontView v{camera, W, H, t->cfg.samples_per_frame, maxDepth, t->rrStart, frameIndex, frameSeed};
	mrad_batch_stats st{};
	MRAD_TRY(wavefrontFrame(&t->wf, t->cfg, t->arena.view, t->range, v, t->output.as<float>(), &st));
	t->last.kernel_nanos = st.kernel_nanos;
	t->last.maximum_physical_kernel_nanos = st.kernel_nanos;
	t->last.physical_launches = (int) std::min<int64_t>(st.steps, 0x7FFFFFFF);
	if (t->cfg.collect_metrics) std::memcpy(t->last.counters, st.counters, sizeof t->last.counters);
	else std::memset(t->last.counters, 0, sizeof t->last.counters);
	const int64_t copyStart = nowNanos();
	if (rgb) MRAD_HIP(hipMemcpy(rgb, t->output.p, (size_t) W * H * 3 * sizeof(float), hipMemcpyDeviceToHost));
	t->last.copy_nanos = rgb ? nowNanos() - copyStart : 0;
	t->last.total_nanos = nowNanos() - frameStart;
	return MRAD_OK;
}

// renderSample (GpuRayTracer.java:317-432).
int mrad_render_sample(mrad_tracer *t, float *rgb, const float camera[12], int maxDepth, int frameIndex,
                       uint64_t frameSeed) {
	clear_error();
	MRAD_TRY(mrad_validate_full_path_depth(maxDepth));
	if (!camera) return set_error(MRAD_ERR_INVALID_ARGUMENT, "GPU scene and camera must not be null");
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	const int W = t->cfg.width, H = t->cfg.height, spp = t->cfg.samples_per_frame, per = t->cfg.samples_per_launch;
	LaunchPlan L;
	if (t->renderForm == MRAD_RENDER_FORM_MEGAKERNEL)
		MRAD_TRY(launchPlan(t, L));                                 // before the device, as the reference orders it
	MRAD_TRY(requireTracer(t));
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident: call mrad_upload_scene first");
	MRAD_HIP(hipSetDevice(t->cfg.device));
	if (t->renderForm == MRAD_RENDER_FORM_WAVEFRONT) return renderWavefront(t, rgb, camera, maxDepth, frameIndex, frameSeed);
	const int64_t frameStart = nowNanos();
	const int launches = (int) L.launches;
	MRAD_TRY(ensure(t->output, (size_t) W * H * 3 * sizeof(float)));
	MRAD_TRY(ensure(t->stats, MRAD_STAT_COUNT * sizeof(unsigned long long)));
	MRAD_TRY(ensureEvents(t, launches - 1));
	const bool metrics = t->cfg.collect_metrics != 0;
	if (metrics) MRAD_HIP(hipMemsetAsync(t->stats.p, 0, MRAD_STAT_COUNT * sizeof(unsigned long long), t->stream));
	RenderParams P{};
	P.S = t->arena.view;
	P.R = t->range;
	std::memcpy(P.cam, camera, sizeof P.cam);
	P.width = W;
	P.height = H;
	P.maxDepth = maxDepth;
	P.rrStart = t->rrStart;
	P.frameIndex = frameIndex;
	P.total = spp;
	P.frameSeed = frameSeed;
	P.out = t->output.as<float>();
	P.stats = metrics ? t->stats.as<unsigned long long>() : nullptr;
	const int stack = t->cfg.bvh_stack_size, lb = t->leafBatch;
	const RenderFn plain = metrics ? renderFor<true, false>(stack, lb) : renderFor<false, false>(stack, lb);
	const RenderFn split = kRenderTargetWaves > 1 ? (metrics ? renderFor<true, true>(stack, lb)
	                                                         : renderFor<false, true>(stack, lb))
	                                              : plain;
	MRAD_HIP(hipEventRecord(t->evStart, t->stream));
	// sample ranges outermost, pixel tiles inside (the reference's order); each
	// pixel keeps its own running sum, so the plan does not change the image
	int l = 0;
	for (int sl = 0; sl < L.sampleLaunches; sl++) {
		P.s0 = sl * per;
		P.s1 = std::min(spp, P.s0 + per);
		for (int ty = 0; ty < L.tileRows; ty++)
			for (int tx = 0; tx < L.tileCols; tx++, l++) {
				P.x0 = tx * L.tileW;
				P.y0 = ty * L.tileH;
				P.regionW = std::min(L.tileW, W - P.x0);
				P.regionH = std::min(L.tileH, H - P.y0);
				P.tilesX = (P.regionW + kTile - 1) / kTile;
				const int tiles = P.tilesX * ((P.regionH + kTile - 1) / kTile);
				P.laneShift = 0;
				while (kRenderBlock == 64 && P.laneShift < 6 && ((int64_t) tiles << P.laneShift) < kRenderTargetWaves &&
				       (2 << P.laneShift) <= P.s1 - P.s0)
					P.laneShift++;
				const int blocks = tiles << P.laneShift;
				hipLaunchKernelGGL(P.laneShift > 0 ? split : plain, dim3(blocks), dim3(kRenderBlock), 0, t->stream, P);
				MRAD_HIP(hipGetLastError());
				if (l + 1 < launches) MRAD_HIP(hipEventRecord(t->boundary[l], t->stream));
			}
	}
	MRAD_HIP(hipEventRecord(t->evEnd, t->stream));
	MRAD_HIP(hipEventSynchronize(t->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, t->evStart, t->evEnd));
	int64_t maxPhys = 0;
	hipEvent_t prev = t->evStart;
	for (int k = 0; k < launches; k++) {
		hipEvent_t end = k + 1 == launches ? t->evEnd : t->boundary[k];
		float pm = 0.0f;
		MRAD_HIP(hipEventElapsedTime(&pm, prev, end));
		const int64_t phys = (int64_t) (pm * 1.0e6f);
		// requireSafePhysicalKernelNanos (GpuRayTracer.java:572-580)
		if (t->cfg.max_physical_kernel_nanos > 0 && phys >= t->cfg.max_physical_kernel_nanos)
			return set_error(MRAD_ERR_ILLEGAL_STATE,
			                 "Physical kernel duration must be greater than zero and less than " +
			                     std::to_string(t->cfg.max_physical_kernel_nanos) + " ns; observed " +
			                     std::to_string(phys) + " ns");
		maxPhys = std::max<int64_t>(maxPhys, phys);
		prev = end;
	}
	t->last.kernel_nanos = (int64_t) (ms * 1.0e6f);
	t->last.maximum_physical_kernel_nanos = std::min(maxPhys, t->last.kernel_nanos);
	t->last.physical_launches = launches;
	const int64_t copyStart = nowNanos();
	if (rgb) MRAD_HIP(hipMemcpy(rgb, t->output.p, (size_t) W * H * 3 * sizeof(float), hipMemcpyDeviceToHost));
	t->last.copy_nanos = rgb ? nowNanos() - copyStart : 0;
	if (metrics) MRAD_HIP(hipMemcpy(t->last.counters, t->stats.p, sizeof t->last.counters, hipMemcpyDeviceToHost));
	else std::memset(t->last.counters, 0, sizeof t->last.counters);
	t->last.total_nanos = nowNanos() - frameStart;
	return MRAD_OK;
}

int mrad_copy_frame(mrad_tracer *t, float *rgb) {
	clear_error();
	MRAD_TRY(requireTracer(t));
	if (!rgb) return set_error(MRAD_ERR_INVALID_ARGUMENT, "RGB output must not be null");
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->output.p) return set_error(MRAD_ERR_ILLEGAL_STATE, "no frame has been rendered");
	MRAD_HIP(hipSetDevice(t->cfg.device));
	MRAD_HIP(hipMemcpy(rgb, t->output.p, (size_t) t->cfg.width * t->cfg.height * 3 * sizeof(float), hipMemcpyDeviceToHost));
	return MRAD_OK;
}

// traceReplay (GpuRayTracer.java:593-716).
int mrad_trace_replay(mrad_tracer *t, const float *rays, int n, int collect, int32_t *flags, float *dist, int32_t *order,
                      int32_t *mat, float *normals, uint32_t *perRay) {
	clear_error();
	MRAD_TRY(replayChecks(t, rays, n, flags, dist, order, mat, normals, perRay));
	std::lock_guard<std::mutex> g(t->mu);
	MRAD_TRY(replayState(t, perRay));
	MRAD_HIP(hipSetDevice(t->cfg.device));
	const size_t N = (size_t) n;
	MRAD_TRY(ensure(t->rays, N * 24));
	MRAD_TRY(ensure(t->flags, N * 4));
	MRAD_TRY(ensure(t->dist, N * 4));
	MRAD_TRY(ensure(t->order, N * 4));
	MRAD_TRY(ensure(t->mat, N * 4));
	MRAD_TRY(ensure(t->normals, N * 12));
	MRAD_TRY(ensure(t->stats, MRAD_STAT_COUNT * 8));
	if (perRay) MRAD_TRY(ensure(t->perRay, N * MRAD_STAT_COUNT * 4));
	ReplayBuffers B;
	B.rays = t->rays.as<float>();
	B.flags = t->flags.as<int>();
	B.dist = t->dist.as<float>();
	B.order = t->order.as<int>();
	B.mat = t->mat.as<int>();
	B.normals = t->normals.as<float>();
	B.perRay = t->perRay.as<uint32_t>();
	B.stats = t->stats.as<unsigned long long>();
	return replayWith(t, B, rays, n, collect, flags, dist, order, mat, normals, perRay, true);
}

// The replay of mrad_trace_replay with its device staging in the library's
// static scratch array instead of buffers it allocates.
int mrad_trace_replay_noalloc(mrad_tracer *t, const float *rays, int n, int collect, int32_t *flags, float *dist,
                              int32_t *order, int32_t *mat, float *normals, uint32_t *perRay) {
	clear_error();
	MRAD_TRY(replayChecks(t, rays, n, flags, dist, order, mat, normals, perRay));
	if (n > MRAD_NOALLOC_MAX_RAYS)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "mrad_trace_replay_noalloc takes at most " +
		                                                std::to_string(MRAD_NOALLOC_MAX_RAYS) + " rays per call, got " +
		                                                std::to_string(n));
	std::lock_guard<std::mutex> g(t->mu);
	MRAD_TRY(replayState(t, perRay));
	MRAD_HIP(hipSetDevice(t->cfg.device));
	unsigned char *base = nullptr;
	MRAD_TRY(scratchFor(t->cfg.device, &base));
	ReplayBuffers B;
	B.stats = reinterpret_cast<unsigned long long *>(base + kScratchStats);
	B.rays = reinterpret_cast<float *>(base + kScratchRays);
	B.flags = reinterpret_cast<int *>(base + kScratchFlags);
	B.dist = reinterpret_cast<float *>(base + kScratchDist);
	B.order = reinterpret_cast<int *>(base + kScratchOrder);
	B.mat = reinterpret_cast<int *>(base + kScratchMat);
	B.normals = reinterpret_cast<float *>(base + kScratchNormals);
	B.perRay = reinterpret_cast<uint32_t *>(base + kScratchPerRay);
	// no per-launch events (creating them is a runtime allocation this entry
	// point avoids): a multi-launch plan reports the whole interval as its
	// longest launch
	return replayWith(t, B, rays, n, collect, flags, dist, order, mat, normals, perRay, false);
}

int mrad_get_frame_stats(mrad_tracer *t, mrad_frame_stats *out) {
	clear_error();
	if (!t || !out) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer and output are required");
	std::lock_guard<std::mutex> g(t->mu);
	*out = t->last;
	return MRAD_OK;
}

int mrad_device_info(mrad_tracer *t, char *name, int nameLen, char *arch, int archLen) {
	clear_error();
	MRAD_TRY(requireTracer(t));
	hipDeviceProp_t p;
	MRAD_HIP(hipGetDeviceProperties(&p, t->cfg.device));
	if (name && nameLen > 0) { std::strncpy(name, p.name, nameLen - 1); name[nameLen - 1] = '\0'; }
	if (arch && archLen > 0) { std::strncpy(arch, p.gcnArchName, archLen - 1); arch[archLen - 1] = '\0'; }
	return MRAD_OK;
}

}  // extern "C"
