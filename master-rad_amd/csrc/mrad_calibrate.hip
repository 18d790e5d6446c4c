// Primitive-cost calibration on gfx950: the measurement kernel of the
// reference's calibration campaign (gpu/PrimitiveCostBenchmark.java:717-873)
// applied to this library's traversal-time operations.
//
// Each thread runs `iterations` iterations; an iteration picks a ray of the
// chosen profile and `copies` records by hash, and calls one operation per
// record through a non-inlined wrapper whose result is consumed in full.
// Operation 0 (setup) returns the wrapper's seed result without any test, so
// operation time minus setup time is the cost of the test itself:
//   1 sphere, 2 box, 3 affine sphere, 4 affine box  the traversal-time tests
//                                                   (distance and accept predicate)
//   5 node AABB                                     one entry distance
//   6 interior traversal                            a node record's two entry
//                                                   distances, ordering and pushes
//   7 plane                                         one plane distance
// The host protocol (pairing, sub-launches, medians) is in mrad/calibrate.py.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "mrad_host.hpp"

namespace mrad {

struct CostData {
	const float *rays;          // profiles x records x 6 (origin, direction)
	const float4 *spheres;      // {centre, radius}
	const float4 *boxes;        // 2 per record: {lo.xyz, hi.x}, {hi.y, hi.z, 0, 0}
	const float4 *affines;      // 3 per record: inverse rows with the centre in .w
	const float *planes;        // wire format, 26 per record
	const float4 *aabbs;        // 2 per record, box layout
	const float4 *nodes;        // 4 per record: two child bounds and {left, right, 0, 0}
	int records;
};

template <int kOp>
__device__ __attribute__((noinline)) Best costOperation(const CostData D, int rec, Ray ray) {
	Best b;
	b.t = 64.0f + fabsf(ray.o.x) + fabsf(ray.o.y) + fabsf(ray.o.z);
	b.rank = (uint32_t) rec;
	b.kind = 0;
	b.index = rec;
	float t;
	if constexpr (kOp == 1) {
		if (testSphere(D.spheres[rec], ray, b.t, t)) { b.t = t; b.kind = 1; }
	} else if constexpr (kOp == 2) {
		if (testBox(D.boxes[2 * rec], D.boxes[2 * rec + 1], ray, b.t, t)) { b.t = t; b.kind = 2; }
	} else if constexpr (kOp == 3) {
		if (testAffineSphere(D.affines + 3 * rec, ray, b.t, t)) { b.t = t; b.kind = 3; }
	} else if constexpr (kOp == 4) {
		if (testAffineBox(D.affines + 3 * rec, ray, b.t, t)) { b.t = t; b.kind = 4; }
	} else if constexpr (kOp == 5) {
		const float4 a = D.aabbs[2 * rec], c = D.aabbs[2 * rec + 1];
		const float e = entryDistance(a.x, a.y, a.z, a.w, c.x, c.y, ray, b.t);
		if (isfinite(e)) { b.t = e; b.kind = 5; }
	} else if constexpr (kOp == 6) {
		const float4 *n = D.nodes + 4 * rec;
		const float4 a = n[0], bb = n[1], cc = n[2];
		const int4 nd = *reinterpret_cast<const int4 *>(n + 3);
		const float le = entryDistance(a.x, a.y, a.z, a.w, bb.x, bb.y, ray, b.t);
		const float re = entryDistance(bb.z, bb.w, cc.x, cc.y, cc.z, cc.w, ray, b.t);
		int nearN = nd.x, farN = nd.y;
		float nearE = le, farE = re;
		if (re < le) { nearN = nd.y; farN = nd.x; nearE = re; farE = le; }
		int node[2];
		float entry[2];
		int size = 0;
		if (isfinite(farE)) { node[size] = farN; entry[size++] = farE; }
		if (isfinite(nearE)) { node[size] = nearN; entry[size++] = nearE; }
		if (size > 0) { b.t = entry[size - 1]; b.index = node[size - 1]; b.kind = 6; }
	} else if constexpr (kOp == 7) {
		const float *p = D.planes + 26 * (size_t) rec;
		const f3 n = mk(p[0], p[1], p[2]);
		const float denom = dot(n, ray.d);
		if (denom != 0.0f) {
			t = -(dot(n, ray.o) + p[3]) / denom;
			if (t > kEps && t < b.t) { b.t = t; b.kind = 7; }
		}
	}
	return b;
}

__device__ __forceinline__ float consume(const Best &b) {
	return b.t * 0.000001f + (float) b.kind * 0.000073f + (float) b.rank * 0.0000001f + (float) b.index * 0.0000003f;
}

template <int kOp>
__device__ __forceinline__ float costLoop(const CostData &D, int tid, int profile, int offset, int iterations,
                                          int copies) {
	float acc = 0.0f;
	const uint32_t mask = (uint32_t) D.records - 1u;
	for (int i = 0; i < iterations; i++) {
		const uint32_t sample = hash_u32((uint32_t) (offset + i) * 0x9E3779B9u + (uint32_t) tid);
		const float *r = D.rays + ((size_t) profile * D.records + (sample & mask)) * 6;
		const f3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
		const Ray ray = makeRay(o, d, true);
		for (int c = 0; c < copies; c++) {
			const uint32_t rs = hash_u32(sample ^ ((uint32_t) (c + 1) * 0x85EBCA6Bu));
			const int rec = (int) (rs & mask);
			const Best b = costOperation<kOp>(D, rec, ray);
			acc += o.x * 0.00013f + o.y * 0.00017f + o.z * 0.00019f + d.x * 0.00023f + d.y * 0.00029f +
			       d.z * 0.00031f + (float) (rec & 7) * 0.0000001f + consume(b);
		}
	}
	return acc;
}

__global__ void costKernel(CostData D, int op, int profile, int offset, int iterations, int copies, float *out) {
	const int tid = blockIdx.x * blockDim.x + threadIdx.x;
	float v;
	switch (op) {
		case 0: v = costLoop<0>(D, tid, profile, offset, iterations, copies); break;
		case 1: v = costLoop<1>(D, tid, profile, offset, iterations, copies); break;
		case 2: v = costLoop<2>(D, tid, profile, offset, iterations, copies); break;
		case 3: v = costLoop<3>(D, tid, profile, offset, iterations, copies); break;
		case 4: v = costLoop<4>(D, tid, profile, offset, iterations, copies); break;
		case 5: v = costLoop<5>(D, tid, profile, offset, iterations, copies); break;
		case 6: v = costLoop<6>(D, tid, profile, offset, iterations, copies); break;
		case 7: v = costLoop<7>(D, tid, profile, offset, iterations, copies); break;
		default: v = 0.0f; break;
	}
	out[tid] = v;
}

}  // namespace mrad

using namespace mrad;

struct mrad_cost {
	int device = 0;
	int records = 0, profiles = 0;
	hipStream_t stream = nullptr;
	hipEvent_t start = nullptr, end = nullptr;
	DevBuf data, out;
	CostData D{};
};

static void cost_free(mrad_cost *c) {
	if (!c) return;
	(void) hipSetDevice(c->device);
	if (c->start) (void) hipEventDestroy(c->start);
	if (c->end) (void) hipEventDestroy(c->end);
	if (c->stream) (void) hipStreamDestroy(c->stream);
	c->data.release();
	c->out.release();
	delete c;
}

static int cost_init(mrad_cost *c, const float *w) {
	const int n = c->records, P = c->profiles;
	const float *rays = w;
	const float *sph = rays + (size_t) P * n * 6;
	const float *box = sph + (size_t) n * 5;
	const float *aff = box + (size_t) n * 7;
	const float *pla = aff + (size_t) n * 22;
	const float *aabb = pla + (size_t) n * 26;
	const float *pair = aabb + (size_t) n * 6;
	// device layout, 16-byte aligned sections
	std::vector<float4> spheres(n), boxes(2 * n), affines(3 * n), aabbs(2 * n), nodes(4 * n);
	for (int i = 0; i < n; i++) {
		const float *s = sph + 5 * (size_t) i, *b = box + 7 * (size_t) i, *a = aff + 22 * (size_t) i;
		const float *q = aabb + 6 * (size_t) i, *p = pair + 12 * (size_t) i;
		spheres[i] = make_float4(s[0], s[1], s[2], s[3]);
		boxes[2 * i] = make_float4(b[0], b[1], b[2], b[3]);
		boxes[2 * i + 1] = make_float4(b[4], b[5], 0.0f, 0.0f);
		for (int r = 0; r < 3; r++) affines[3 * i + r] = make_float4(a[4 * r], a[4 * r + 1], a[4 * r + 2], a[4 * r + 3]);
		aabbs[2 * i] = make_float4(q[0], q[1], q[2], q[3]);
		aabbs[2 * i + 1] = make_float4(q[4], q[5], 0.0f, 0.0f);
		nodes[4 * i] = make_float4(p[0], p[1], p[2], p[3]);
		nodes[4 * i + 1] = make_float4(p[4], p[5], p[6], p[7]);
		nodes[4 * i + 2] = make_float4(p[8], p[9], p[10], p[11]);
		int4 nd = make_int4(0, 1, 0, 0);
		std::memcpy(&nodes[4 * i + 3], &nd, sizeof nd);
	}
	const size_t rayBytes = ((size_t) P * n * 6 * sizeof(float) + 255) & ~(size_t) 255;
	const size_t f4 = sizeof(float4);
	const size_t planeBytes = ((size_t) n * 26 * sizeof(float) + 255) & ~(size_t) 255;
	const size_t total = rayBytes + (size_t) n * f4 * (1 + 2 + 3 + 2 + 4) + planeBytes;
	MRAD_TRY(ensure(c->data, total));
	char *base = c->data.as<char>();
	size_t off = 0;
	auto put = [&](const void *src, size_t bytes, size_t padded) -> char * {
		char *dst = base + off;
		hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
		off += padded;
		return e == hipSuccess ? dst : nullptr;
	};
	char *pr = put(rays, (size_t) P * n * 6 * sizeof(float), rayBytes);
	char *ps = put(spheres.data(), n * f4, n * f4);
	char *pb = put(boxes.data(), 2 * n * f4, 2 * n * f4);
	char *pa = put(affines.data(), 3 * n * f4, 3 * n * f4);
	char *pq = put(aabbs.data(), 2 * n * f4, 2 * n * f4);
	char *pn = put(nodes.data(), 4 * n * f4, 4 * n * f4);
	char *pp = put(pla, (size_t) n * 26 * sizeof(float), planeBytes);
	if (!pr || !ps || !pb || !pa || !pq || !pn || !pp)
		return set_error(MRAD_ERR_HIP, "hipMemcpy of the calibration workload failed");
	c->D = CostData{reinterpret_cast<const float *>(pr), reinterpret_cast<const float4 *>(ps),
	                reinterpret_cast<const float4 *>(pb), reinterpret_cast<const float4 *>(pa),
	                reinterpret_cast<const float *>(pp), reinterpret_cast<const float4 *>(pq),
	                reinterpret_cast<const float4 *>(pn), n};
	return MRAD_OK;
}

extern "C" int mrad_cost_create(int device, const float *workload, int records, int profiles, int *compute_units,
                                mrad_cost **out) {
	clear_error();
	if (!out || !workload || !compute_units || records <= 0 || (records & (records - 1)) != 0 || profiles <= 0)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "calibration workload: records must be a power of two, "
		                                            "profiles positive, pointers non-null");
	*out = nullptr;
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return set_error(MRAD_ERR_UNAVAILABLE, "no HIP device");
	if (device < 0 || device >= count) return set_error(MRAD_ERR_INVALID_ARGUMENT, "device index out of range");
	mrad_cost *c = new mrad_cost();
	c->device = device;
	c->records = records;
	c->profiles = profiles;
	int rc = MRAD_OK;
	if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
	    hipEventCreate(&c->start) != hipSuccess || hipEventCreate(&c->end) != hipSuccess)
		rc = set_error(MRAD_ERR_HIP, "calibration stream or events");
	if (rc == MRAD_OK && hipDeviceGetAttribute(compute_units, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
		rc = set_error(MRAD_ERR_HIP, "compute-unit count of the calibration device");
	if (rc == MRAD_OK) rc = cost_init(c, workload);
	if (rc != MRAD_OK) {
		cost_free(c);
		return rc;
	}
	*out = c;
	return MRAD_OK;
}

extern "C" int mrad_cost_time(mrad_cost *c, int operation, int profile, int iteration_offset, int iterations,
                              int copies, int blocks, int threads, float *elapsed_ms) {
	clear_error();
	if (!c || !elapsed_ms || operation < 0 || operation > 7 || profile < 0 || profile >= c->profiles ||
	    iterations <= 0 || copies <= 0 || blocks <= 0 || threads <= 0 || threads > 1024 || threads % 64 != 0)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "calibration launch: operation 0..7, profile in range, "
		                                            "positive counts, threads a multiple of 64 up to 1024");
	MRAD_HIP(hipSetDevice(c->device));
	MRAD_TRY(ensure(c->out, (size_t) blocks * threads * sizeof(float)));
	MRAD_HIP(hipEventRecord(c->start, c->stream));
	hipLaunchKernelGGL(costKernel, dim3(blocks), dim3(threads), 0, c->stream, c->D, operation, profile,
	                   iteration_offset, iterations, copies, c->out.as<float>());
	MRAD_HIP(hipGetLastError());
	MRAD_HIP(hipEventRecord(c->end, c->stream));
	MRAD_HIP(hipEventSynchronize(c->end));
	MRAD_HIP(hipEventElapsedTime(elapsed_ms, c->start, c->end));
	return MRAD_OK;
}

// The per-thread accumulators of the last mrad_cost_time launch (blocks x
// threads floats, thread order).
extern "C" int mrad_cost_output(mrad_cost *c, float *out, int n) {
	clear_error();
	if (!c || !out || n <= 0) return set_error(MRAD_ERR_INVALID_ARGUMENT, "calibration output: handle, buffer, count");
	if ((size_t) n * sizeof(float) > c->out.cap || !c->out.p)
		return set_error(MRAD_ERR_INVALID_ARGUMENT, "calibration output: more values than the last launch wrote");
	MRAD_HIP(hipSetDevice(c->device));
	MRAD_HIP(hipMemcpy(out, c->out.p, (size_t) n * sizeof(float), hipMemcpyDeviceToHost));
	return MRAD_OK;
}

extern "C" void mrad_cost_destroy(mrad_cost *c) { cost_free(c); }
