// Kernels and host runtime behind include/mrad.h: the MI355X replacement of
// the reference's JCuda/NVRTC launcher (src/xyz/marsavic/gfxlab/gpu/GpuRayTracer.java)
// and of its two kernels, renderKernel and traceReplayKernel
// (gpu/GpuKernelSources.java:1167-1393).  Code objects are compiled ahead of
// time for gfx950; there is no runtime compilation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mrad.h"
#include "mrad_device.hpp"
#include "mrad_error.hpp"

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

constexpr int kBlock = 256;      // 4 wave64s, one 16x16 pixel tile
constexpr int kTile = 16;
constexpr int kWaves = kBlock / 64;

// Per-block LDS for the short traversal stacks of its 4 waves.
struct StackLds {
	int node[kWaves * (kShortStack > 0 ? kShortStack : 1) * 64];
	float entry[kWaves * (kShortStack > 0 ? kShortStack : 1) * 64];
};
__device__ __forceinline__ LdsStack laneStack(StackLds &s) {
	const int base = (threadIdx.x >> 6) * kShortStack * 64 + (threadIdx.x & 63);
	return {s.node + base, s.entry + base};
}

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
	float cam[12];
	int width, height, maxDepth, rrStart, frameIndex, s0, s1, total, tilesX;
	uint64_t frameSeed;
	float *out;
	unsigned long long *stats;
};

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
// (GpuKernelSources.java:1295-1310).
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

// renderKernel (GpuKernelSources.java:1167-1315).  One thread per pixel runs
// samples [s0, s1) of the frame and keeps the running sum in registers; the
// float sum order and the last-sample normalization are the reference's.
// A wave covers an 8x8 pixel block for primary-ray coherence.
template <bool kMetrics, int kCap>
__global__ __launch_bounds__(kBlock) void renderKernel(RenderParams P) {
	const DevScene &S = P.S;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = xcdContiguous(blockIdx.x, gridDim.x);
	const int tileX = tile % P.tilesX, tileY = tile / P.tilesX;
	const int x = tileX * kTile + (wave & 1) * 8 + (lane & 7);
	const int y = tileY * kTile + (wave >> 1) * 8 + (lane >> 3);
	const bool active = x < P.width && y < P.height;
	__shared__ StackLds stackLds;
	const LdsStack lds = laneStack(stackLds);
	Counters<kMetrics> c;
	c.zero();
	if (active) {
		const uint32_t pixel = (uint32_t) y * (uint32_t) P.width + (uint32_t) x;
		const size_t idx = (size_t) pixel * 3;
		const f3 camPos = mk(P.cam[0], P.cam[1], P.cam[2]), camF = mk(P.cam[3], P.cam[4], P.cam[5]);
		const f3 camR = mk(P.cam[6], P.cam[7], P.cam[8]), camU = mk(P.cam[9], P.cam[10], P.cam[11]);
		const float aspect = (float) P.width / (float) P.height;
		f3 acc = P.s0 > 0 ? mk(P.out[idx], P.out[idx + 1], P.out[idx + 2]) : mk(0.0f, 0.0f, 0.0f);
		for (int s = P.s0; s < P.s1; s++) {
			uint32_t rng = path_seed(P.frameSeed, P.frameIndex, pixel, s);
			const float jx = rand01(rng) - 0.5f;
			const float jy = rand01(rng) - 0.5f;
			const float u = ((((float) x + 0.5f) + jx) * (2.0f / (float) P.width) - 1.0f) * aspect;
			const float v = 1.0f - ((((float) y + 0.5f) + jy) * (2.0f / (float) P.height));
			const f3 dir = camF + camR * u + camU * v;
			const f3 col = tracePath<kMetrics, kCap>(S, camPos, dir, P.maxDepth, P.rrStart, rng, c, lds);
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
	const float *rays;
	int n;
	int *flags;
	float *dist;
	int *order;
	int *mat;
	float *normals;
	uint32_t *perRay;
	unsigned long long *stats;
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
		const bool found = traceClosest<kMetrics, kCap>(P.S, o, d, best, c, lds);
		c.add(0, 1);
		HitDetail h;
		if (found) h = resolveHit(P.S, o, d, best);
		P.flags[r] = found ? 1 : 0;
		P.dist[r] = found ? best.t : INFINITY;
		P.order[r] = found ? best.order : -1;
		P.mat[r] = found ? h.matIndex : -1;
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
template <bool M>
RenderFn renderFor(int cap) {
	if (cap <= 64) return renderKernel<M, 64>;
	return renderKernel<M, 256>;
}
template <bool M>
ReplayFn replayFor(int cap) {
	if (cap <= 64) return replayKernel<M, 64>;
	return replayKernel<M, 256>;
}

// ---------------------------------------------------------------- runtime

struct DevBuf {
	void *p = nullptr;
	size_t cap = 0;
	void release() {
		if (p) (void) hipFree(p);
		p = nullptr;
		cap = 0;
	}
};

#define MRAD_HIP(expr)                                                                         \
	do {                                                                                       \
		hipError_t e_ = (expr);                                                                \
		if (e_ != hipSuccess)                                                                  \
			return set_error(MRAD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
	} while (0)

int ensure(DevBuf &b, size_t bytes) {
	if (bytes == 0) bytes = 16;
	if (b.cap >= bytes) return MRAD_OK;
	b.release();
	hipError_t e = hipMalloc(&b.p, bytes);
	if (e != hipSuccess) {
		b.p = nullptr;
		return set_error(MRAD_ERR_OUT_OF_MEMORY, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
	}
	b.cap = bytes;
	return MRAD_OK;
}

}  // namespace mrad

using namespace mrad;

struct mrad_tracer {
	std::mutex mu;
	mrad_tracer_config cfg{};
	bool available = false;
	std::string unavailableReason;
	hipStream_t stream = nullptr;
	hipEvent_t evStart = nullptr, evEnd = nullptr;
	std::vector<hipEvent_t> boundary;
	// resident scene
	DevBuf nodes, refs, spheres, sphereMat, boxes, affS, affSN, affB, affBN, planes, materials;
	DevScene scene{};
	bool hasScene = false;
	uint64_t residentKey = 0, residentRevision = 0;
	// frame and replay buffers
	DevBuf output, stats, rays, flags, dist, order, mat, normals, perRay;
	int rrStart = -1;
	mrad_frame_stats last{};
};

namespace {

int64_t nowNanos() {
	return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int requireTracer(mrad_tracer *t) {
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	if (!t->available)
		return set_error(MRAD_ERR_UNAVAILABLE, "HIP path tracer is unavailable: " + t->unavailableReason);
	return MRAD_OK;
}

template <typename T>
int upload(hipStream_t s, DevBuf &b, const std::vector<T> &host) {
	int rc = ensure(b, host.size() * sizeof(T));
	if (rc != MRAD_OK) return rc;
	if (!host.empty()) MRAD_HIP(hipMemcpyAsync(b.p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, s));
	return MRAD_OK;
}

inline float4 f4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
inline float ibits(int v) { float f; std::memcpy(&f, &v, 4); return f; }

// uploadScene (GpuRayTracer.java:873-932): wire arrays re-laid-out into the
// device records described in mrad_device.hpp.
int uploadScene(mrad_tracer *t, const mrad_scene_arrays *a) {
	const int nc = a->node_count;
	if (a->sphere_count < 0 || a->box_count < 0 || a->plane_count < 0 || a->affine_sphere_count < 0 ||
	    a->affine_box_count < 0 || nc < 0 || a->primitive_ref_count < 0 || a->root_count < 0 || a->material_count < 0)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "scene counts must be non-negative");
	if ((a->sphere_count && !a->spheres) || (a->box_count && !a->boxes) || (a->plane_count && !a->planes) ||
	    (a->affine_sphere_count && !a->affine_spheres) || (a->affine_box_count && !a->affine_boxes) ||
	    (nc && (!a->node_bounds || !a->node_data)) || (a->primitive_ref_count && !a->primitive_refs) ||
	    (a->root_count && !a->root_indices) || (a->material_count && !a->materials))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "scene arrays must be non-null for non-zero counts");
	if (a->root_count > 4) return set_error(MRAD_ERR_INVALID_ARGUMENT, "at most 4 BVH roots are supported");
	std::vector<float4> nodes((size_t) nc * 4);
	for (int n = 0; n < nc; n++) {
		const int32_t *nd = a->node_data + 4 * (size_t) n;
		float4 *r = nodes.data() + 4 * (size_t) n;
		if (nd[0] >= 0) {
			if (nd[0] >= nc || nd[1] < 0 || nd[1] >= nc)
				return set_error(MRAD_ERR_INVALID_ARGUMENT, "BVH child index out of range at node " + std::to_string(n));
			const float *L = a->node_bounds + 6 * (size_t) nd[0];
			const float *R = a->node_bounds + 6 * (size_t) nd[1];
			r[0] = f4(L[0], L[1], L[2], L[3]);
			r[1] = f4(L[4], L[5], R[0], R[1]);
			r[2] = f4(R[2], R[3], R[4], R[5]);
		} else {
			r[0] = r[1] = r[2] = f4(0, 0, 0, 0);
		}
		r[3] = f4(ibits(nd[0]), ibits(nd[1]), ibits(nd[2]), ibits(nd[3]));
	}
	std::vector<float4> sph(a->sphere_count);
	std::vector<int> sphMat(a->sphere_count);
	for (int i = 0; i < a->sphere_count; i++) {
		const float *s = a->spheres + 5 * (size_t) i;
		sph[i] = f4(s[0], s[1], s[2], s[3]);
		sphMat[i] = (int) s[4];
	}
	std::vector<float4> box((size_t) a->box_count * 2);
	for (int i = 0; i < a->box_count; i++) {
		const float *b = a->boxes + 7 * (size_t) i;
		box[2 * i] = f4(b[0], b[1], b[2], b[3]);
		box[2 * i + 1] = f4(b[4], b[5], ibits((int) b[6]), 0.0f);
	}
	auto affine = [](const float *src, int n, std::vector<float4> &test, std::vector<float4> &nrm) {
		test.resize((size_t) n * 3);
		nrm.resize((size_t) n * 3);
		for (int i = 0; i < n; i++) {
			const float *d = src + 22 * (size_t) i;
			test[3 * i] = f4(d[0], d[1], d[2], d[3]);
			test[3 * i + 1] = f4(d[4], d[5], d[6], d[7]);
			test[3 * i + 2] = f4(d[8], d[9], d[10], d[11]);
			nrm[3 * i] = f4(d[12], d[13], d[14], d[15]);
			nrm[3 * i + 1] = f4(d[16], d[17], d[18], d[19]);
			nrm[3 * i + 2] = f4(d[20], ibits((int) d[21]), 0.0f, 0.0f);
		}
	};
	std::vector<float4> aS, aSN, aB, aBN;
	affine(a->affine_spheres, a->affine_sphere_count, aS, aSN);
	affine(a->affine_boxes, a->affine_box_count, aB, aBN);
	std::vector<float> planes(a->planes ? a->planes : nullptr, a->planes ? a->planes + 26 * (size_t) a->plane_count : nullptr);
	std::vector<float4> mats((size_t) a->material_count * 4);
	for (int i = 0; i < a->material_count; i++) {
		const float *m = a->materials + 13 * (size_t) i;
		mats[4 * i] = f4(m[0], m[1], m[2], 0.0f);
		mats[4 * i + 1] = f4(m[3], m[4], m[5], 0.0f);
		mats[4 * i + 2] = f4(m[6], m[7], m[8], 0.0f);
		mats[4 * i + 3] = f4(m[9], m[10], m[11], m[12]);
	}
	std::vector<uint32_t> refs(a->primitive_refs ? (const uint32_t *) a->primitive_refs : nullptr,
	                           a->primitive_refs ? (const uint32_t *) a->primitive_refs + a->primitive_ref_count : nullptr);
	int rc;
	hipStream_t s = t->stream;
	if ((rc = upload(s, t->nodes, nodes)) || (rc = upload(s, t->refs, refs)) || (rc = upload(s, t->spheres, sph)) ||
	    (rc = upload(s, t->sphereMat, sphMat)) || (rc = upload(s, t->boxes, box)) || (rc = upload(s, t->affS, aS)) ||
	    (rc = upload(s, t->affSN, aSN)) || (rc = upload(s, t->affB, aB)) || (rc = upload(s, t->affBN, aBN)) ||
	    (rc = upload(s, t->planes, planes)) || (rc = upload(s, t->materials, mats)))
		return rc;
	MRAD_HIP(hipStreamSynchronize(s));
	DevScene &S = t->scene;
	S = DevScene{};
	S.nodes = (const float4 *) t->nodes.p;
	S.refs = (const uint32_t *) t->refs.p;
	S.spheres = (const float4 *) t->spheres.p;
	S.sphereMat = (const int *) t->sphereMat.p;
	S.boxes = (const float4 *) t->boxes.p;
	S.affS = (const float4 *) t->affS.p;
	S.affSN = (const float4 *) t->affSN.p;
	S.affB = (const float4 *) t->affB.p;
	S.affBN = (const float4 *) t->affBN.p;
	S.planes = (const float *) t->planes.p;
	S.materials = (const float4 *) t->materials.p;
	S.sphereCount = a->sphere_count;
	S.boxCount = a->box_count;
	S.affSCount = a->affine_sphere_count;
	S.affBCount = a->affine_box_count;
	S.planeCount = a->plane_count;
	S.nodeCount = t->cfg.bvh_traversal ? nc : 0;
	S.refCount = t->cfg.bvh_traversal ? a->primitive_ref_count : 0;
	S.rootCount = t->cfg.bvh_traversal ? a->root_count : 0;
	S.useBvh = t->cfg.bvh_traversal ? 1 : 0;
	S.materialCount = a->material_count;
	for (int r = 0; r < a->root_count && r < 4; r++) {
		int ri = a->root_indices[r];
		S.roots[r] = ri;
		for (int k = 0; k < 6; k++) S.rootBounds[r][k] = (ri >= 0 && ri < nc) ? a->node_bounds[6 * (size_t) ri + k] : 0.0f;
	}
	for (int k = 0; k < 3; k++) S.bg[k] = a->background[k];
	S.stackLimit = t->cfg.bvh_stack_size;
	t->hasScene = true;
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
	mrad_tracer *t = new (std::nothrow) mrad_tracer();
	if (!t) return set_error(MRAD_ERR_OUT_OF_MEMORY, "tracer allocation failed");
	t->cfg = c;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count <= 0) {
		t->unavailableReason = e != hipSuccess ? hipGetErrorString(e) : "no HIP device";
	} else if (c.device < 0 || c.device >= count) {
		t->unavailableReason = "device ordinal out of range";
	} else if ((e = hipSetDevice(c.device)) != hipSuccess || (e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking)) != hipSuccess ||
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
			for (DevBuf *b : {&t->nodes, &t->refs, &t->spheres, &t->sphereMat, &t->boxes, &t->affS, &t->affSN, &t->affB,
			                  &t->affBN, &t->planes, &t->materials, &t->output, &t->stats, &t->rays, &t->flags, &t->dist,
			                  &t->order, &t->mat, &t->normals, &t->perRay})
				b->release();
			for (hipEvent_t e : t->boundary) (void) hipEventDestroy(e);
			if (t->evStart) (void) hipEventDestroy(t->evStart);
			if (t->evEnd) (void) hipEventDestroy(t->evEnd);
			if (t->stream) (void) hipStreamDestroy(t->stream);
		}
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

int mrad_upload_scene(mrad_tracer *t, const mrad_scene_arrays *a, uint64_t key) {
	clear_error();
	int rc = requireTracer(t);
	if (rc) return rc;
	if (!a) return set_error(MRAD_ERR_INVALID_ARGUMENT, "GPU scene must not be null");
	std::lock_guard<std::mutex> g(t->mu);
	MRAD_HIP(hipSetDevice(t->cfg.device));
	if (t->hasScene && t->residentKey == key && t->residentRevision == a->revision) {
		t->last.upload_nanos = 0;
		return MRAD_OK;
	}
	int64_t t0 = nowNanos();
	t->hasScene = false;
	rc = uploadScene(t, a);
	if (rc) return rc;
	t->residentKey = key;
	t->residentRevision = a->revision;
	t->last.upload_nanos = std::max<int64_t>(1, nowNanos() - t0);
	return MRAD_OK;
}

// renderSample (GpuRayTracer.java:317-432).
int mrad_render_sample(mrad_tracer *t, float *rgb, const float camera[12], int maxDepth, int frameIndex, uint64_t frameSeed) {
	clear_error();
	int rc = mrad_validate_full_path_depth(maxDepth);
	if (rc) return rc;
	if (!camera) return set_error(MRAD_ERR_INVALID_ARGUMENT, "GPU scene and camera must not be null");
	rc = requireTracer(t);
	if (rc) return rc;
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident: call mrad_upload_scene first");
	MRAD_HIP(hipSetDevice(t->cfg.device));
	const int64_t frameStart = nowNanos();
	const int W = t->cfg.width, H = t->cfg.height, spp = t->cfg.samples_per_frame, per = t->cfg.samples_per_launch;
	const int launches = (spp + per - 1) / per;
	if ((rc = ensure(t->output, (size_t) W * H * 3 * sizeof(float)))) return rc;
	if ((rc = ensure(t->stats, MRAD_STAT_COUNT * sizeof(unsigned long long)))) return rc;
	if ((rc = ensureEvents(t, launches - 1))) return rc;
	const bool metrics = t->cfg.collect_metrics != 0;
	if (metrics) MRAD_HIP(hipMemsetAsync(t->stats.p, 0, MRAD_STAT_COUNT * sizeof(unsigned long long), t->stream));
	RenderParams P{};
	P.S = t->scene;
	std::memcpy(P.cam, camera, sizeof P.cam);
	P.width = W;
	P.height = H;
	P.maxDepth = maxDepth;
	P.rrStart = t->rrStart;
	P.frameIndex = frameIndex;
	P.total = spp;
	P.tilesX = (W + kTile - 1) / kTile;
	P.frameSeed = frameSeed;
	P.out = (float *) t->output.p;
	P.stats = metrics ? (unsigned long long *) t->stats.p : nullptr;
	const int blocks = P.tilesX * ((H + kTile - 1) / kTile);
	RenderFn fn = metrics ? renderFor<true>(t->cfg.bvh_stack_size) : renderFor<false>(t->cfg.bvh_stack_size);
	MRAD_HIP(hipEventRecord(t->evStart, t->stream));
	for (int l = 0; l < launches; l++) {
		P.s0 = l * per;
		P.s1 = std::min(spp, P.s0 + per);
		hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, t->stream, P);
		MRAD_HIP(hipGetLastError());
		if (l + 1 < launches) MRAD_HIP(hipEventRecord(t->boundary[l], t->stream));
	}
	MRAD_HIP(hipEventRecord(t->evEnd, t->stream));
	MRAD_HIP(hipEventSynchronize(t->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, t->evStart, t->evEnd));
	int64_t maxPhys = 0;
	hipEvent_t prev = t->evStart;
	for (int l = 0; l < launches; l++) {
		hipEvent_t end = l + 1 == launches ? t->evEnd : t->boundary[l];
		float pm = 0.0f;
		MRAD_HIP(hipEventElapsedTime(&pm, prev, end));
		maxPhys = std::max<int64_t>(maxPhys, (int64_t) (pm * 1.0e6f));
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
	int rc = requireTracer(t);
	if (rc) return rc;
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
	if (!rays || n <= 0) return set_error(MRAD_ERR_INVALID_ARGUMENT, "Replay rays must contain origin.xyz and direction.xyz.");
	if (!flags || !dist || !order || !mat || !normals) return set_error(MRAD_ERR_INVALID_ARGUMENT, "replay outputs must not be null");
	int rc = requireTracer(t);
	if (rc) return rc;
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident: call mrad_upload_scene first");
	if (perRay && !t->cfg.collect_metrics)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "per-ray counters need a tracer created with collect_metrics");
	MRAD_HIP(hipSetDevice(t->cfg.device));
	const int64_t t0 = nowNanos();
	const size_t N = (size_t) n;
	if ((rc = ensure(t->rays, N * 24)) || (rc = ensure(t->flags, N * 4)) || (rc = ensure(t->dist, N * 4)) ||
	    (rc = ensure(t->order, N * 4)) || (rc = ensure(t->mat, N * 4)) || (rc = ensure(t->normals, N * 12)) ||
	    (rc = ensure(t->stats, MRAD_STAT_COUNT * 8)))
		return rc;
	const bool metrics = collect && t->cfg.collect_metrics;
	if (perRay && (rc = ensure(t->perRay, N * MRAD_STAT_COUNT * 4))) return rc;
	MRAD_HIP(hipMemcpyAsync(t->rays.p, rays, N * 24, hipMemcpyHostToDevice, t->stream));
	MRAD_HIP(hipMemsetAsync(t->stats.p, 0, MRAD_STAT_COUNT * 8, t->stream));
	ReplayParams P{};
	P.S = t->scene;
	P.rays = (const float *) t->rays.p;
	P.n = n;
	P.flags = (int *) t->flags.p;
	P.dist = (float *) t->dist.p;
	P.order = (int *) t->order.p;
	P.mat = (int *) t->mat.p;
	P.normals = (float *) t->normals.p;
	P.perRay = perRay ? (uint32_t *) t->perRay.p : nullptr;
	P.stats = (metrics || perRay) ? (unsigned long long *) t->stats.p : nullptr;
	const bool instrumented = metrics || perRay;
	ReplayFn fn = instrumented ? replayFor<true>(t->cfg.bvh_stack_size) : replayFor<false>(t->cfg.bvh_stack_size);
	MRAD_HIP(hipEventRecord(t->evStart, t->stream));
	hipLaunchKernelGGL(fn, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, t->stream, P);
	MRAD_HIP(hipGetLastError());
	MRAD_HIP(hipEventRecord(t->evEnd, t->stream));
	MRAD_HIP(hipEventSynchronize(t->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, t->evStart, t->evEnd));
	const int64_t c0 = nowNanos();
	MRAD_HIP(hipMemcpy(flags, t->flags.p, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(dist, t->dist.p, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(order, t->order.p, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(mat, t->mat.p, N * 4, hipMemcpyDeviceToHost));
	MRAD_HIP(hipMemcpy(normals, t->normals.p, N * 12, hipMemcpyDeviceToHost));
	if (perRay) MRAD_HIP(hipMemcpy(perRay, t->perRay.p, N * MRAD_STAT_COUNT * 4, hipMemcpyDeviceToHost));
	if (instrumented) MRAD_HIP(hipMemcpy(t->last.counters, t->stats.p, sizeof t->last.counters, hipMemcpyDeviceToHost));
	else std::memset(t->last.counters, 0, sizeof t->last.counters);
	t->last.copy_nanos = nowNanos() - c0;
	t->last.kernel_nanos = (int64_t) (ms * 1.0e6f);
	t->last.maximum_physical_kernel_nanos = t->last.kernel_nanos;
	t->last.physical_launches = 1;
	t->last.total_nanos = nowNanos() - t0;
	return MRAD_OK;
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
	int rc = requireTracer(t);
	if (rc) return rc;
	hipDeviceProp_t p;
	MRAD_HIP(hipGetDeviceProperties(&p, t->cfg.device));
	if (name && nameLen > 0) { std::strncpy(name, p.name, nameLen - 1); name[nameLen - 1] = '\0'; }
	if (arch && archLen > 0) { std::strncpy(arch, p.gcnArchName, archLen - 1); arch[archLen - 1] = '\0'; }
	return MRAD_OK;
}

}  // extern "C"
