// The tracer handle behind mrad_create/mrad_destroy, shared by the render,
// replay and adjoint translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "mrad_host.hpp"

struct mrad_tracer {
	std::mutex mu;
	mrad_tracer_config cfg{};
	bool available = false;
	std::string unavailableReason;
	hipStream_t stream = nullptr;
	hipEvent_t evStart = nullptr, evEnd = nullptr;
	std::vector<hipEvent_t> boundary;
	mrad::DevArena arena;
	mrad::SceneRange range{};
	int materialCount = 0;
	bool hasScene = false;
	uint64_t residentKey = 0, residentRevision = 0;
	mrad::DevBuf output, stats, rays, flags, dist, order, mat, normals, perRay;
	mrad::DevBuf gradImage, gradAlbedo, gradEmission;
	int rrStart = -1;
	int leafBatch = 0;                  // megakernel leaf batching: 0 or kLeafBatchLanes
	mrad_frame_stats last{};
	int renderForm = MRAD_RENDER_FORM_MEGAKERNEL;
	mrad_batch *wf = nullptr;           // the wavefront form's batch, created on first use
};

namespace mrad {

// One frame of the resident scene in the wavefront form (mrad_wavefront.hip):
// a batch of one view over the tracer's own device arena, created on first
// use and kept for later frames; the image goes to `out` (W x H x 3 device
// floats), bit-identical to the megakernel's.
struct WavefrontView {
	const float *cam;
	int width, height, spp, maxDepth, rrStart, frameIndex;
	uint64_t frameSeed;
};
int wavefrontFrame(mrad_batch **batch, const mrad_tracer_config &cfg, const DevScene &S, const SceneRange &R,
                   const WavefrontView &v, float *out, mrad_batch_stats *stats);

inline int requireTracer(mrad_tracer *t) {
	if (!t) return set_error(MRAD_ERR_INVALID_ARGUMENT, "tracer handle is null");
	if (!t->available)
		return set_error(MRAD_ERR_UNAVAILABLE, "HIP path tracer is unavailable: " + t->unavailableReason);
	return MRAD_OK;
}

inline int primitiveCount(const SceneRange &R) { return R.sphCount + R.boxCount + R.asCount + R.abCount + R.planeCount; }

}  // namespace mrad
