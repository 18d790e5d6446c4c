// Path-replay adjoint (no counterpart in the reference): gradients of an image
// loss with respect to per-primitive diffuse albedo and emittance, after
// Vicini, Speierer and Jakob, "Path Replay Backpropagation" (SIGGRAPH 2021).
//
// The forward pass is mrad_render_sample.  The backward pass re-traces every
// path from its recorded seed with the forward arithmetic, so it visits the
// same vertices without storing any path between the passes.  Within one
// thread the vertices of the current path (throughput, lobe weight, diffuse
// factor, emission, primitive) are kept in private memory and swept backwards:
//   dLoss/dLe_i     += delta * T_i
//   dLoss/dalbedo_i += delta * T_i * g_i * Lin_i,     g_i = 1/(p_lobe q_rr) for a diffuse lobe, else 0
//   Lin_{i-1}        = Le_i + w_i * Lin_i
// where delta is dLoss/dpixel / spp and Lin of the last vertex is the
// background when its scattered ray missed, else zero.  The sweep needs no
// division, so zero albedos are exact.  Lobe-selection and roulette
// probabilities are treated as constants (their decisions are detached from
// the parameters), which matches central differences with common random
// numbers for Lambertian materials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "mrad_host.hpp"
#include "mrad_kernels_common.hpp"
#include "mrad_tracer.hpp"

namespace mrad {

// Block of the backward kernel: MRAD_ADJOINT_BLOCK threads in 8x8-pixel
// waves, 256 (a 16x16 tile of 4 waves) or 64 (one 8x8 wave per block).
#ifndef MRAD_ADJOINT_BLOCK
#define MRAD_ADJOINT_BLOCK 64
#endif
constexpr int kAdjBlock = MRAD_ADJOINT_BLOCK;
static_assert(kAdjBlock == 64 || kAdjBlock == 256, "adjoint block is one or four waves");
constexpr int kAdjTile = kAdjBlock == 256 ? 16 : 8;
constexpr int kMaxVertices = 32;
// The samples of a tile are split over `groups` blocks when the frame has too
// few tiles to fill the chip: gradients are FP64 atomic sums, whose order is
// free, so a pixel's samples need not run in one wave.  groups is the smallest
// count that gives MRAD_ADJOINT_TARGET_WAVES waves (at most one per sample).
// 65,536 waves: 16 groups of 2 samples for the 4,096 tiles of a 512x512, 32-spp
// frame (backward 7.2 -> 3.7 ms; 2 / 4 / 8 groups: 5.3 / 4.2 / 3.8 ms), 3 groups
// at 1080p, one for frames of 65,536 tiles or more.
#ifndef MRAD_ADJOINT_TARGET_WAVES
#define MRAD_ADJOINT_TARGET_WAVES 65536
#endif
constexpr int kAdjTargetWaves = MRAD_ADJOINT_TARGET_WAVES;

struct AdjointParams {
	DevScene S;
	SceneRange R;
	float cam[12];
	int width, height, spp, maxDepth, rrStart, frameIndex, tilesX;
	int groups;              // blocks per tile, each taking a contiguous share of the samples
	uint64_t frameSeed;
	const float *gradImage;
	double *gAlbedo;         // FP64 sums: the per-vertex terms change sign with the
	double *gEmission;       // pixel residual, so FP32 sums lose the small net values
};

// Occupancy the backward kernel is compiled for (waves per SIMD; 0: the
// compiler's choice).  Left to itself the compiler keeps part of the vertex
// records in registers (147 VGPRs, 3 waves per SIMD); a target of 5 puts them
// in scratch and gives 82 VGPRs, 6 waves per SIMD, and an 18 % shorter pass.
#ifndef MRAD_ADJOINT_WAVES_PER_EU
#define MRAD_ADJOINT_WAVES_PER_EU 5
#endif
#if MRAD_ADJOINT_WAVES_PER_EU > 0
#define MRAD_ADJOINT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(MRAD_ADJOINT_WAVES_PER_EU)))
#else
#define MRAD_ADJOINT_OCCUPANCY
#endif

template <int kCap>
__global__ __launch_bounds__(kAdjBlock) MRAD_ADJOINT_OCCUPANCY void adjointKernel(AdjointParams P) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = blockIdx.x / P.groups, group = blockIdx.x % P.groups;
	const int tileX = tile % P.tilesX, tileY = tile / P.tilesX;
	const int sBegin = (int) ((int64_t) group * P.spp / P.groups), sEnd = (int) ((int64_t) (group + 1) * P.spp / P.groups);
	const int x = tileX * kAdjTile + (wave & 1) * 8 + (lane & 7);
	const int y = tileY * kAdjTile + (wave >> 1) * 8 + (lane >> 3);
	__shared__ StackLdsT<kShortStack, kAdjBlock / 64> stackLds;
	const LdsStack lds = laneStack(stackLds);
	if (x >= P.width || y >= P.height) return;
	const uint32_t pixel = (uint32_t) y * (uint32_t) P.width + (uint32_t) x;
	const float inv = 1.0f / (float) P.spp;
	const f3 delta = mk(P.gradImage[3 * (size_t) pixel] * inv, P.gradImage[3 * (size_t) pixel + 1] * inv,
	                    P.gradImage[3 * (size_t) pixel + 2] * inv);
	if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) return;
	const f3 bg = mk(P.R.bg[0], P.R.bg[1], P.R.bg[2]);
	f3 Tv[kMaxVertices], Wv[kMaxVertices], Lev[kMaxVertices];
	float gv[kMaxVertices];
	int prim[kMaxVertices];
	Counters<false> none;
	for (int s = sBegin; s < sEnd; s++) {
		uint32_t rng = path_seed(P.frameSeed, P.frameIndex, pixel, s);
		f3 o, d;
		cameraRay(P.cam, P.width, P.height, x, y, rng, o, d);
		f3 T = mk(1.0f, 1.0f, 1.0f);
		f3 tail = mk(0.0f, 0.0f, 0.0f);
		int n = 0;
		for (int depth = 0; depth < P.maxDepth; ++depth) {
			Best best;
			if (!traceClosest<false, kCap>(P.S, P.R, o, d, best, none, lds)) {
				tail = bg;
				break;
			}
			const HitDetail h = resolveHit(P.S, o, d, best.t, best.kind, best.index);
			const MaterialF m = loadMaterial(P.S, h.matIndex);
			Tv[n] = T;
			Lev[n] = m.emittance;
			prim[n] = localOrder(P.R, best);
			Wv[n] = mk(0.0f, 0.0f, 0.0f);
			gv[n] = 0.0f;
			n++;
			// scatter (GpuKernelSources.java:1120-1161) keeping the lobe and its probability
			const f3 in = -d;
			const float wD = luminance(m.diffuse), wR = luminance(m.reflective), wT = luminance(m.refractive);
			const float sum = wD + wR + wT;
			if (sum <= 0.0f) break;
			const float xi = rand01(rng) * sum;
			f3 w;
			float p;
			bool diffuse = false;
			if (xi < wD) {
				p = wD / sum;
				diffuse = true;
				d = sampleUnitSphere(rng) + h.normal;
				w = divs(m.diffuse, p);
			} else if (xi < wD + wR) {
				p = wR / sum;
				d = reflectDir(h.normal, in);
				w = divs(m.reflective, p);
			} else {
				p = wT / sum;
				d = refractDir(in, h.normal, m.ior);
				w = divs(m.refractive, p);
			}
			o = h.pos + d * kEps;
			T = T * w;
			const float scale = diffuse ? 1.0f / p : 0.0f;
			float q = 1.0f;
			if (P.rrStart >= 0 && depth + 1 >= P.rrStart && depth + 1 < P.maxDepth) {
				const float mx = fmaxf(T.x, fmaxf(T.y, T.z));
				const float qf = fminf(fmaxf(mx, 0.05f), 0.95f);
				if (rand01(rng) >= qf) {
					Wv[n - 1] = w;
					gv[n - 1] = scale;
					break;
				}
				T = T * (1.0f / qf);
				q = qf;
			}
			Wv[n - 1] = mk(w.x / q, w.y / q, w.z / q);
			gv[n - 1] = scale / q;
			if (depth + 1 == P.maxDepth) break;
		}
		f3 lin = tail;
		for (int i = n - 1; i >= 0; i--) {
			const f3 t = Tv[i];
			const int base = 3 * prim[i];
			const f3 ge = delta * t;
			const f3 ga = ge * lin * gv[i];
			atomicAdd(P.gEmission + base, (double) ge.x);
			atomicAdd(P.gEmission + base + 1, (double) ge.y);
			atomicAdd(P.gEmission + base + 2, (double) ge.z);
			if (gv[i] != 0.0f) {
				atomicAdd(P.gAlbedo + base, (double) ga.x);
				atomicAdd(P.gAlbedo + base + 1, (double) ga.y);
				atomicAdd(P.gAlbedo + base + 2, (double) ga.z);
			}
			lin = Lev[i] + Wv[i] * lin;
		}
	}
}

}  // namespace mrad

using namespace mrad;

extern "C" {

// Overwrites the resident material table (same count), e.g. for the next
// forward pass of an optimisation; it stays until another scene or revision
// is uploaded.  No counterpart in the reference.
int mrad_update_materials(mrad_tracer *t, const float *materials13, int count) {
	clear_error();
	MRAD_TRY(requireTracer(t));
	if (!materials13 || count < 0) return set_error(MRAD_ERR_INVALID_ARGUMENT, "materials must not be null");
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident");
	if (count != t->materialCount)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "material count differs from the resident scene's");
	std::vector<float4> mats((size_t) count * 4);
	for (int i = 0; i < count; i++) {
		const float *m = materials13 + 13 * (size_t) i;
		mats[4 * i] = f4(m[0], m[1], m[2], 0.0f);
		mats[4 * i + 1] = f4(m[3], m[4], m[5], 0.0f);
		mats[4 * i + 2] = f4(m[6], m[7], m[8], 0.0f);
		mats[4 * i + 3] = f4(m[9], m[10], m[11], m[12]);
	}
	MRAD_HIP(hipSetDevice(t->cfg.device));
	MRAD_TRY(upload(t->stream, t->arena.materials, mats));
	MRAD_HIP(hipStreamSynchronize(t->stream));
	t->arena.view.materials = t->arena.materials.as<const float4>();
	return MRAD_OK;
}

// Backward pass of renderSample for the resident scene and this tracer's
// width, height, samples and roulette setting.
int mrad_backward(mrad_tracer *t, const float *gradImage, const float camera[12], int maxDepth, int frameIndex,
                  uint64_t frameSeed, float *gradAlbedo, float *gradEmission, int primCount) {
	clear_error();
	MRAD_TRY(mrad_validate_full_path_depth(maxDepth));
	if (!gradImage || !camera || !gradAlbedo || !gradEmission)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "gradient image, camera and outputs must not be null");
	MRAD_TRY(requireTracer(t));
	std::lock_guard<std::mutex> g(t->mu);
	if (!t->hasScene) return set_error(MRAD_ERR_ILLEGAL_STATE, "no scene is resident: call mrad_upload_scene first");
	if (primCount != primitiveCount(t->range))
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "primitive count differs from the resident scene's");
	MRAD_HIP(hipSetDevice(t->cfg.device));
	const int W = t->cfg.width, H = t->cfg.height;
	const size_t px = (size_t) W * H * 3, np = (size_t) primCount * 3;
	MRAD_TRY(ensure(t->gradImage, px * 4));
	MRAD_TRY(ensure(t->gradAlbedo, np * 8));
	MRAD_TRY(ensure(t->gradEmission, np * 8));
	MRAD_HIP(hipMemcpyAsync(t->gradImage.p, gradImage, px * 4, hipMemcpyHostToDevice, t->stream));
	MRAD_HIP(hipMemsetAsync(t->gradAlbedo.p, 0, np * 8, t->stream));
	MRAD_HIP(hipMemsetAsync(t->gradEmission.p, 0, np * 8, t->stream));
	AdjointParams P{};
	P.S = t->arena.view;
	P.R = t->range;
	std::memcpy(P.cam, camera, sizeof P.cam);
	P.width = W;
	P.height = H;
	P.spp = t->cfg.samples_per_frame;
	P.maxDepth = maxDepth;
	P.rrStart = t->rrStart;
	P.frameIndex = frameIndex;
	P.frameSeed = frameSeed;
	P.tilesX = (W + kAdjTile - 1) / kAdjTile;
	P.gradImage = t->gradImage.as<const float>();
	P.gAlbedo = t->gradAlbedo.as<double>();
	P.gEmission = t->gradEmission.as<double>();
	const int tiles = P.tilesX * ((H + kAdjTile - 1) / kAdjTile);
	const int wavesPerTile = kAdjBlock / 64;
	P.groups = std::max(1, std::min(P.spp, (kAdjTargetWaves + tiles * wavesPerTile - 1) / (tiles * wavesPerTile)));
	const int blocks = tiles * P.groups;
	const int64_t t0 = nowNanos();
	MRAD_HIP(hipEventRecord(t->evStart, t->stream));
	if (t->cfg.bvh_stack_size <= 64) hipLaunchKernelGGL(adjointKernel<64>, dim3(blocks), dim3(kAdjBlock), 0, t->stream, P);
	else hipLaunchKernelGGL(adjointKernel<256>, dim3(blocks), dim3(kAdjBlock), 0, t->stream, P);
	MRAD_HIP(hipGetLastError());
	MRAD_HIP(hipEventRecord(t->evEnd, t->stream));
	MRAD_HIP(hipEventSynchronize(t->evEnd));
	float ms = 0.0f;
	MRAD_HIP(hipEventElapsedTime(&ms, t->evStart, t->evEnd));
	std::vector<double> host(np);
	MRAD_HIP(hipMemcpy(host.data(), t->gradAlbedo.p, np * 8, hipMemcpyDeviceToHost));
	for (size_t k = 0; k < np; k++) gradAlbedo[k] = (float) host[k];
	MRAD_HIP(hipMemcpy(host.data(), t->gradEmission.p, np * 8, hipMemcpyDeviceToHost));
	for (size_t k = 0; k < np; k++) gradEmission[k] = (float) host[k];
	t->last.kernel_nanos = (int64_t) (ms * 1.0e6f);
	t->last.maximum_physical_kernel_nanos = t->last.kernel_nanos;
	t->last.physical_launches = 1;
	t->last.copy_nanos = 0;
	std::memset(t->last.counters, 0, sizeof t->last.counters);
	t->last.total_nanos = nowNanos() - t0;
	return MRAD_OK;
}

}  // extern "C"
