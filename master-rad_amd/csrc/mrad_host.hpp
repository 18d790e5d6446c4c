// Host-side helpers shared by the tracer, batch and adjoint runtimes:
// device buffers, HIP status mapping, and the arena builder that turns the
// reference's wire arrays (GpuScene.java:24-31) into the device records of
// mrad_device.hpp, rebasing indices so many scenes can share one arena.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mrad.h"
#include "mrad_device.hpp"
#include "mrad_error.hpp"

// 1 stores the two children of each BVH node side by side in the arena (see
// nodeSlots); 0 keeps the scene's node order.
#ifndef MRAD_NODE_PAIRS
#define MRAD_NODE_PAIRS 1
#endif
// Levels below the roots laid out breadth first at the start of a scene's
// node range (the part kernels may stage in LDS); deeper nodes follow depth
// first.
#ifndef MRAD_TOP_LEVELS
#define MRAD_TOP_LEVELS 5
#endif

namespace mrad {

#define MRAD_HIP(expr)                                                                                 \
	do {                                                                                               \
		hipError_t e_ = (expr);                                                                        \
		if (e_ != hipSuccess)                                                                          \
			return ::mrad::set_error(MRAD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
	} while (0)

#define MRAD_TRY(expr)           \
	do {                         \
		int rc_ = (expr);        \
		if (rc_ != MRAD_OK)      \
			return rc_;          \
	} while (0)

struct DevBuf {
	void *p = nullptr;
	size_t cap = 0;
	void release() {
		if (p) (void) hipFree(p);
		p = nullptr;
		cap = 0;
	}
	template <typename T>
	T *as() const { return static_cast<T *>(p); }
};

inline int ensure(DevBuf &b, size_t bytes) {
	if (bytes == 0) bytes = 16;
	if (b.cap >= bytes) return MRAD_OK;
	b.release();
	hipError_t e = hipMalloc(&b.p, bytes);
	if (e != hipSuccess) {
		b.p = nullptr;
		return set_error(MRAD_ERR_OUT_OF_MEMORY,
		                 "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
	}
	b.cap = bytes;
	return MRAD_OK;
}

template <typename T>
inline int upload(hipStream_t s, DevBuf &b, const std::vector<T> &host) {
	MRAD_TRY(ensure(b, host.size() * sizeof(T)));
	if (!host.empty()) MRAD_HIP(hipMemcpyAsync(b.p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, s));
	return MRAD_OK;
}

inline int64_t nowNanos() {
	return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
	    .count();
}

inline float4 f4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
inline float ibits(int v) { float f; std::memcpy(&f, &v, 4); return f; }

// Host staging of an arena: concatenated device records of one or more scenes.
struct HostArena {
	std::vector<float4> nodes, spheres, boxes, affS, affSN, affB, affBN, materials;
	std::vector<int> sphereMat;
	std::vector<uint32_t> refs;
	std::vector<float> planes;

	// Whether every node and box bound is 0 or of magnitude in [2^-35, 2^58],
	// the condition under which the device takes slab quotients from the
	// ray's reciprocals (see the ray constants in mrad_device.hpp).
	static bool slabBoundsQualify(const mrad_scene_arrays *a) {
		auto ok = [](float v) {
			const float m = std::fabs(v);
			return v == 0.0f || (m >= 0x1p-35f && m <= 0x1p58f);
		};
		for (size_t i = 0; i < 6 * (size_t) a->node_count; i++)
			if (!ok(a->node_bounds[i])) return false;
		for (int i = 0; i < a->box_count; i++)
			for (int k = 0; k < 6; k++)
				if (!ok(a->boxes[7 * (size_t) i + k])) return false;
		return true;
	}

	// Appends one scene; fills its SceneRange.  useBvh and stackLimit come
	// from the caller's configuration.  A scene that fails validation leaves
	// the arena as it was.
	int append(const mrad_scene_arrays *a, bool useBvh, int stackLimit, SceneRange &R) {
		const size_t n0 = nodes.size(), s0 = spheres.size(), b0 = boxes.size(), as0 = affS.size(),
		             asn0 = affSN.size(), ab0 = affB.size(), abn0 = affBN.size(), m0 = materials.size(),
		             sm0 = sphereMat.size(), r0 = refs.size(), p0 = planes.size();
		const int rc = appendScene(a, useBvh, stackLimit, R);
		if (rc != MRAD_OK) {
			nodes.resize(n0); spheres.resize(s0); boxes.resize(b0); affS.resize(as0); affSN.resize(asn0);
			affB.resize(ab0); affBN.resize(abn0); materials.resize(m0); sphereMat.resize(sm0); refs.resize(r0);
			planes.resize(p0);
		}
		return rc;
	}

  private:
	// Arena slot of every node of a scene's BVH, relative to its (even) base;
	// returns the number of slots and sets `top` to the number of slots laid
	// out breadth first.  With MRAD_NODE_PAIRS the two children of each
	// interior node take slots 2k and 2k + 1, so they share one 128-byte line:
	// lanes of a wave that part at a node fetch one line, not two.  The roots
	// come first (each alone at an even slot, one pad slot after it), then the
	// child pairs of the first MRAD_TOP_LEVELS levels breadth first, so that
	// the first slots of a scene hold exactly its top levels; below them the
	// pairs are laid out depth first.  Storage order is invisible to
	// traversal: children, bounds and visiting order are the scene's own.  A
	// BVH that is not a forest of trees over all its nodes (a node reached
	// twice, a child index out of range, a node no root reaches) keeps the
	// scene's order, and the validation that follows reports any bad index.
	static int nodeSlots(const mrad_scene_arrays *a, std::vector<int> &slot, int &top) {
		const int nc = a->node_count;
		slot.assign((size_t) nc, -1);
		top = 0;
		int next = 0;
		bool ok = MRAD_NODE_PAIRS != 0;
		auto pair = [&](int n, std::vector<int> &out) {
			const int32_t *nd = a->node_data + 4 * (size_t) n;
			if (nd[0] < 0) return;
			const int l = nd[0], rr = nd[1];
			if (l >= nc || rr < 0 || rr >= nc || l == rr || slot[l] >= 0 || slot[rr] >= 0) { ok = false; return; }
			slot[l] = next;
			slot[rr] = next + 1;
			next += 2;
			out.push_back(l);
			out.push_back(rr);
		};
		std::vector<int> level, below;
		for (int r = 0; ok && r < a->root_count; r++) {
			const int ri = a->root_indices[r];
			if (ri < 0 || ri >= nc) continue;
			if (slot[ri] >= 0) { ok = false; break; }
			slot[ri] = next;
			next += 2;
			level.push_back(ri);
		}
		for (int d = 0; ok && d < MRAD_TOP_LEVELS && !level.empty(); d++) {
			below.clear();
			for (int n : level) {
				pair(n, below);
				if (!ok) break;
			}
			level.swap(below);
		}
		top = next;
		std::vector<int> todo(level.rbegin(), level.rend());
		while (ok && !todo.empty()) {
			const int n = todo.back();
			todo.pop_back();
			below.clear();
			pair(n, below);
			if (!below.empty()) {
				todo.push_back(below[1]);
				todo.push_back(below[0]);
			}
		}
		for (int n = 0; ok && n < nc; n++)
			if (slot[n] < 0) ok = false;
		if (!ok) {
			for (int n = 0; n < nc; n++) slot[n] = n;
			top = 0;
			return nc;
		}
		return next;
	}

	int appendScene(const mrad_scene_arrays *a, bool useBvh, int stackLimit, SceneRange &R) {
		const int nc = a->node_count;
		if (a->sphere_count < 0 || a->box_count < 0 || a->plane_count < 0 || a->affine_sphere_count < 0 ||
		    a->affine_box_count < 0 || nc < 0 || a->primitive_ref_count < 0 || a->root_count < 0 ||
		    a->material_count < 0)
			return set_error(MRAD_ERR_INVALID_ARGUMENT, "scene counts must be non-negative");
		if ((a->sphere_count && !a->spheres) || (a->box_count && !a->boxes) || (a->plane_count && !a->planes) ||
		    (a->affine_sphere_count && !a->affine_spheres) || (a->affine_box_count && !a->affine_boxes) ||
		    (nc && (!a->node_bounds || !a->node_data)) || (a->primitive_ref_count && !a->primitive_refs) ||
		    (a->root_count && !a->root_indices) || (a->material_count && !a->materials))
			return set_error(MRAD_ERR_INVALID_ARGUMENT, "scene arrays must be non-null for non-zero counts");
		if (a->root_count > 4) return set_error(MRAD_ERR_INVALID_ARGUMENT, "at most 4 BVH roots are supported");
		// node slots: an even base, so that sibling pairs fill 128-byte lines
		if (nodes.size() / 4 % 2) nodes.resize(nodes.size() + 4, f4(0, 0, 0, 0));
		std::vector<int> slot;
		int topSlots = 0;
		const int slots = nodeSlots(a, slot, topSlots);
		const size_t nodeBase = nodes.size() / 4, refBase = refs.size(), sphBase = spheres.size(),
		             boxBase = boxes.size() / 2, asBase = affS.size() / 3, abBase = affB.size() / 3,
		             planeBase = planes.size() / 26, matBase = materials.size() / 4;
		const size_t sizes[] = {nodeBase + slots, refBase + a->primitive_ref_count, sphBase + a->sphere_count,
		                        boxBase + a->box_count, asBase + a->affine_sphere_count, abBase + a->affine_box_count,
		                        matBase + a->material_count};
		for (size_t s : sizes)
			if (s > kRefIndexMask) return set_error(MRAD_ERR_INVALID_ARGUMENT, "arena exceeds 2^30 records");
		const int mb = (int) matBase;
		auto mat = [mb](int m) { return m < 0 ? m : m + mb; };
		// Material indices of the wire records (a float holding an integer):
		// -1 (no material, GpuScene.materialIndexOf) or an index into this
		// scene's table.  Anything else, non-finite values included, would
		// make every kernel read past the table (or another scene's), so the
		// whole scene is rejected before any value is converted.
		const int mc = a->material_count;
		auto matOk = [mc](float v) { return v == -1.0f || (v >= 0.0f && (double) v < (double) mc && v == std::floor(v)); };
		auto badMat = [](const char *what, int i) {
			return set_error(MRAD_ERR_INVALID_ARGUMENT,
			                 std::string(what) + " " + std::to_string(i) + ": material index out of range");
		};
		for (int i = 0; i < a->sphere_count; i++)
			if (!matOk(a->spheres[5 * (size_t) i + 4])) return badMat("sphere", i);
		for (int i = 0; i < a->box_count; i++)
			if (!matOk(a->boxes[7 * (size_t) i + 6])) return badMat("box", i);
		for (int i = 0; i < a->affine_sphere_count; i++)
			if (!matOk(a->affine_spheres[22 * (size_t) i + 21])) return badMat("affine sphere", i);
		for (int i = 0; i < a->affine_box_count; i++)
			if (!matOk(a->affine_boxes[22 * (size_t) i + 21])) return badMat("affine box", i);
		for (int i = 0; i < a->plane_count; i++)
			for (int k = 23; k < 26; k++)
				if (!matOk(a->planes[26 * (size_t) i + k])) return badMat("plane", i);
		for (int i = 0; i < a->primitive_ref_count; i++) {
			const uint32_t ref = (uint32_t) a->primitive_refs[i];
			const uint32_t type = ref & kRefTypeMask;
			const size_t base = type == kRefSphere ? sphBase : type == kRefBox ? boxBase : type == kRefAffineSphere ? asBase : abBase;
			const int count = type == kRefSphere ? a->sphere_count : type == kRefBox ? a->box_count
			                : type == kRefAffineSphere ? a->affine_sphere_count : a->affine_box_count;
			if ((ref & kRefIndexMask) >= (uint32_t) count)
				return set_error(MRAD_ERR_INVALID_ARGUMENT, "primitive ref " + std::to_string(i) + " out of range");
			const size_t idx = (ref & kRefIndexMask) + base;
			refs.push_back(type | (uint32_t) idx);
		}
		for (int i = 0; i < a->sphere_count; i++) {
			const float *s = a->spheres + 5 * (size_t) i;
			spheres.push_back(f4(s[0], s[1], s[2], s[3]));
			sphereMat.push_back(mat((int) s[4]));
		}
		for (int i = 0; i < a->box_count; i++) {
			const float *b = a->boxes + 7 * (size_t) i;
			boxes.push_back(f4(b[0], b[1], b[2], b[3]));
			boxes.push_back(f4(b[4], b[5], ibits(mat((int) b[6])), 0.0f));
		}
		auto affine = [&mat](const float *src, int n, std::vector<float4> &test, std::vector<float4> &nrm) {
			for (int i = 0; i < n; i++) {
				const float *d = src + 22 * (size_t) i;
				test.push_back(f4(d[0], d[1], d[2], d[3]));
				test.push_back(f4(d[4], d[5], d[6], d[7]));
				test.push_back(f4(d[8], d[9], d[10], d[11]));
				nrm.push_back(f4(d[12], d[13], d[14], d[15]));
				nrm.push_back(f4(d[16], d[17], d[18], d[19]));
				nrm.push_back(f4(d[20], ibits(mat((int) d[21])), 0.0f, 0.0f));
			}
		};
		affine(a->affine_spheres, a->affine_sphere_count, affS, affSN);
		affine(a->affine_boxes, a->affine_box_count, affB, affBN);
		for (int i = 0; i < a->plane_count; i++) {
			const float *p = a->planes + 26 * (size_t) i;
			size_t o = planes.size();
			planes.insert(planes.end(), p, p + 26);
			for (int k = 5; k < 8; k++) {
				const int m = (int) p[18 + k];
				if (m >= 0 && m + mb >= (1 << 24))
					return set_error(MRAD_ERR_INVALID_ARGUMENT, "plane material index exceeds float precision");
				planes[o + 18 + k] = (float) mat(m);
			}
		}
		for (int i = 0; i < a->material_count; i++) {
			const float *m = a->materials + 13 * (size_t) i;
			materials.push_back(f4(m[0], m[1], m[2], 0.0f));
			materials.push_back(f4(m[3], m[4], m[5], 0.0f));
			materials.push_back(f4(m[6], m[7], m[8], 0.0f));
			materials.push_back(f4(m[9], m[10], m[11], m[12]));
		}
		// nodes: child pairs, indices rebased.  A leaf of one ref carries that
		// primitive's test record in the 48 bytes an interior node uses for its
		// children's bounds, and the rebased ref in .y, so testing it needs no
		// fetch beyond the node record itself.
		nodes.resize(nodes.size() + (size_t) slots * 4, f4(0, 0, 0, 0));
		for (int n = 0; n < nc; n++) {
			const int32_t *nd = a->node_data + 4 * (size_t) n;
			float4 *r = nodes.data() + 4 * (nodeBase + slot[n]);
			if (nd[0] >= 0) {
				if (nd[0] >= nc || nd[1] < 0 || nd[1] >= nc)
					return set_error(MRAD_ERR_INVALID_ARGUMENT, "BVH child index out of range at node " + std::to_string(n));
				const float *L = a->node_bounds + 6 * (size_t) nd[0];
				const float *Rb = a->node_bounds + 6 * (size_t) nd[1];
				r[0] = f4(L[0], L[1], L[2], L[3]);
				r[1] = f4(L[4], L[5], Rb[0], Rb[1]);
				r[2] = f4(Rb[2], Rb[3], Rb[4], Rb[5]);
				r[3] = f4(ibits(slot[nd[0]] + (int) nodeBase), ibits(slot[nd[1]] + (int) nodeBase), ibits(nd[2]),
				          ibits(nd[3]));
			} else {
				if (nd[3] < 0 || nd[2] < 0 || (int64_t) nd[2] + nd[3] > a->primitive_ref_count)
					return set_error(MRAD_ERR_INVALID_ARGUMENT, "BVH leaf refs out of range at node " + std::to_string(n));
				r[0] = r[1] = r[2] = f4(0, 0, 0, 0);
				int y = nd[1];
				if (nd[3] == 1) {
					const uint32_t ref = refs[refBase + nd[2]];
					const uint32_t idx = ref & kRefIndexMask, type = ref & kRefTypeMask;
					const float4 *src = type == kRefSphere ? &spheres[idx] : type == kRefBox ? &boxes[2 * (size_t) idx]
					                  : type == kRefAffineSphere ? &affS[3 * (size_t) idx] : &affB[3 * (size_t) idx];
					const int words = type == kRefSphere ? 1 : type == kRefBox ? 2 : 3;
					for (int k = 0; k < words; k++) r[k] = src[k];
					y = (int) ref;
				}
				r[3] = f4(ibits(nd[0]), ibits(y), ibits(nd[2] + (int) refBase), ibits(nd[3]));
			}
		}
		R = SceneRange{};
		R.rootCount = a->root_count;
		for (int r = 0; r < a->root_count; r++) {
			const int ri = a->root_indices[r];
			const bool ok = ri >= 0 && ri < nc;
			R.roots[r] = ok ? slot[ri] + (int) nodeBase : -1;
			for (int k = 0; k < 6; k++) R.rootBounds[r][k] = ok ? a->node_bounds[6 * (size_t) ri + k] : 0.0f;
		}
		R.nodeBase = (int) nodeBase;
		R.topSlots = topSlots;
		R.refBase = (int) refBase;
		R.refCount = a->primitive_ref_count;
		R.sphBase = (int) sphBase; R.sphCount = a->sphere_count;
		R.boxBase = (int) boxBase; R.boxCount = a->box_count;
		R.asBase = (int) asBase; R.asCount = a->affine_sphere_count;
		R.abBase = (int) abBase; R.abCount = a->affine_box_count;
		R.planeBase = (int) planeBase; R.planeCount = a->plane_count;
		R.useBvh = (useBvh && nc > 0) ? 1 : 0;
		R.stackLimit = stackLimit;
		R.fastSlabs = slabBoundsQualify(a) ? 1 : 0;
		for (int k = 0; k < 3; k++) R.bg[k] = a->background[k];
		return MRAD_OK;
	}
};

// Device copy of an arena.
struct DevArena {
	DevBuf nodes, refs, spheres, sphereMat, boxes, affS, affSN, affB, affBN, planes, materials;
	DevScene view{};
	int upload(hipStream_t s, const HostArena &h) {
		MRAD_TRY(mrad::upload(s, nodes, h.nodes));
		MRAD_TRY(mrad::upload(s, refs, h.refs));
		MRAD_TRY(mrad::upload(s, spheres, h.spheres));
		MRAD_TRY(mrad::upload(s, sphereMat, h.sphereMat));
		MRAD_TRY(mrad::upload(s, boxes, h.boxes));
		MRAD_TRY(mrad::upload(s, affS, h.affS));
		MRAD_TRY(mrad::upload(s, affSN, h.affSN));
		MRAD_TRY(mrad::upload(s, affB, h.affB));
		MRAD_TRY(mrad::upload(s, affBN, h.affBN));
		MRAD_TRY(mrad::upload(s, planes, h.planes));
		MRAD_TRY(mrad::upload(s, materials, h.materials));
		MRAD_HIP(hipStreamSynchronize(s));
		view.nodes = nodes.as<const float4>();
		view.refs = refs.as<const uint32_t>();
		view.spheres = spheres.as<const float4>();
		view.sphereMat = sphereMat.as<const int>();
		view.boxes = boxes.as<const float4>();
		view.affS = affS.as<const float4>();
		view.affSN = affSN.as<const float4>();
		view.affB = affB.as<const float4>();
		view.affBN = affBN.as<const float4>();
		view.planes = planes.as<const float>();
		view.materials = materials.as<const float4>();
		return MRAD_OK;
	}
	void release() {
		for (DevBuf *b : {&nodes, &refs, &spheres, &sphereMat, &boxes, &affS, &affSN, &affB, &affBN, &planes, &materials})
			b->release();
	}
};

}  // namespace mrad
