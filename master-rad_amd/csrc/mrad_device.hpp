// Device-side arithmetic of the path-tracing hot path for gfx950.
//
// Semantics follow src/xyz/marsavic/gfxlab/gpu/GpuKernelSources.java; every
// expression keeps the reference's operation order, and the library is built
// with -ffp-contract=off so that only the fmaf calls the reference writes are
// fused.  What differs from the reference is data layout and work placement:
//
//   * BVH nodes are 64-byte records: a parent holds both children's bounds
//     (48 B) and its own {left|kind, right, first, count}, so one interior
//     expansion is one aligned 64-byte fetch instead of 16 + 2 x 24 bytes in
//     three places.
//   * During traversal a primitive test computes only its candidate distance
//     and the reference's "found" predicate.  Normal and material of the
//     closest hit are recomputed once, after traversal, by the same
//     arithmetic on the winning primitive, so the per-test fetch of an affine
//     primitive is 48 bytes (inverse rows with the centre) instead of 88.
//   * All scenes of a batch share one arena: node, ref and primitive arrays
//     are concatenated and their indices rebased on upload; a SceneRange
//     names one scene's slices.  Within a scene, primitive order is compared
//     through a rank (type << 30 | arena index), which orders exactly like the
//     reference's global order (spheres, boxes, affine spheres, affine
//     boxes); the reference's order number is rebuilt from it on output.
//
// Accept rules are the reference's: BVH leaves test against
// nextafterf(best.t, +inf) and accept on (t, order) (GpuKernelSources.java:772-806);
// the linear scan and planes test against best.t itself (:338-734, :1015-1018).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrad {

constexpr float kEps = 1e-4f;
constexpr uint32_t kRefTypeMask = 0xC0000000u;
constexpr uint32_t kRefIndexMask = 0x3FFFFFFFu;
constexpr uint32_t kRefSphere = 0x00000000u;
constexpr uint32_t kRefBox = 0x40000000u;
constexpr uint32_t kRefAffineSphere = 0x80000000u;
constexpr uint32_t kRefAffineBox = 0xC0000000u;
constexpr int kKindSphere = 0, kKindBox = 1, kKindAffineSphere = 2, kKindAffineBox = 3, kKindPlane = 4;
constexpr int kStatCount = 14;

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 v, float s) { return {v.x * s, v.y * s, v.z * s}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// float3 / s is v * (1/s) in the reference (GpuKernelSources.java:87-90).
__device__ __forceinline__ f3 divs(f3 v, float s) { float inv = 1.0f / s; return {v.x * inv, v.y * inv, v.z * inv}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 normalizeSafe(f3 v) { float len = sqrtf(fmaxf(dot(v, v), 1e-20f)); return divs(v, len); }
__device__ __forceinline__ float comp(f3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// GpuKernelSources.java:101-110
__device__ inline f3 normalizeTransformedNormal(f3 v) {
	float l2 = dot(v, v);
	if (l2 >= 1.0e-20f && isfinite(l2)) return divs(v, sqrtf(l2));
	float scale = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
	if (!(scale > 0.0f) || !isfinite(scale)) return mk(0.0f, 0.0f, 0.0f);
	f3 s = mk(v.x / scale, v.y / scale, v.z / scale);
	return divs(s, sqrtf(dot(s, s)));
}

// triple32 hash and rand01 (GpuKernelSources.java:165-179).
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
	x ^= x >> 17; x *= 0xed5ad4bbu;
	x ^= x >> 11; x *= 0xac4c1b51u;
	x ^= x >> 15; x *= 0x31848babu;
	x ^= x >> 14;
	return x;
}
__device__ __forceinline__ float rand01(uint32_t &s) {
	s = hash_u32(s + 0x9e3779b9u);
	return (float) (s >> 8) * 0x1p-24f;
}
// GpuKernelSources.java:181-183
__device__ __forceinline__ float luminance(f3 c) { return c.x * 0.212655f + c.y * 0.715158f + c.z * 0.072187f; }

// Single-precision sine and cosine after the Cephes library (S. L. Moshier,
// sinf.c/cosf.c): octant reduction by a three-part pi/4 and its degree-7/8
// minimax polynomials, in plain float arithmetic so host and device agree
// bit for bit.  Arguments here lie in [0, 2 pi).
__device__ inline void sincos_cephes(float xx, float &sOut, float &cOut) {
	const float FOPI = 1.27323954473516f;
	const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
	float x = fabsf(xx);
	int signS = xx < 0.0f ? -1 : 1, signC = 1;
	int j = (int) (x * FOPI);
	float y = (float) j;
	if (j & 1) { j += 1; y += 1.0f; }
	j &= 7;
	if (j > 3) { j -= 4; signS = -signS; signC = -signC; }
	if (j > 1) signC = -signC;
	x = ((x - y * DP1) - y * DP2) - y * DP3;
	float z = x * x;
	float pc = 2.443315711809948e-5f * z - 1.388731625493765e-3f;
	pc = pc * z + 4.166664568298827e-2f;
	pc = pc * z * z - 0.5f * z + 1.0f;
	float ps = -1.9515295891e-4f * z + 8.3321608736e-3f;
	ps = ps * z - 1.6666654611e-1f;
	ps = ps * z * x + x;
	bool swap = (j == 1 || j == 2);
	float s = swap ? pc : ps;
	float c = swap ? ps : pc;
	sOut = signS < 0 ? -s : s;
	cOut = signC < 0 ? -c : c;
}

// Per-path seed (GpuKernelSources.java:1231-1253).
__device__ __forceinline__ uint32_t path_seed(uint64_t frameSeed, int frameIndex, uint32_t pixel, int sampleOffset) {
	uint32_t base = hash_u32((uint32_t) (frameSeed & 0xFFFFFFFFull));
	base ^= hash_u32((uint32_t) (frameSeed >> 32));
	base ^= hash_u32((uint32_t) frameIndex);
	base ^= hash_u32(pixel);
	base = hash_u32(base);
	uint32_t seed = base ^ hash_u32((uint32_t) sampleOffset);
	seed ^= hash_u32((uint32_t) sampleOffset * 0x9E3779B9u);
	return hash_u32(seed);
}

// Camera ray of renderKernel (GpuKernelSources.java:1256-1264).
__device__ __forceinline__ void cameraRay(const float *cam, int width, int height, int x, int y, uint32_t &rng, f3 &o,
                                          f3 &d) {
	const float jx = rand01(rng) - 0.5f;
	const float jy = rand01(rng) - 0.5f;
	const float aspect = (float) width / (float) height;
	const float u = ((((float) x + 0.5f) + jx) * (2.0f / (float) width) - 1.0f) * aspect;
	const float v = 1.0f - ((((float) y + 0.5f) + jy) * (2.0f / (float) height));
	o = mk(cam[0], cam[1], cam[2]);
	d = mk(cam[3], cam[4], cam[5]) + mk(cam[6], cam[7], cam[8]) * u + mk(cam[9], cam[10], cam[11]) * v;
}

// ------------------------------------------------------------ device arena

struct MaterialF { f3 diffuse, reflective, refractive, emittance; float ior; };

// Global arrays of the arena (one scene or a whole batch).
struct DevScene {
	const float4 *nodes;        // 4 x float4 per node (see header comment)
	const uint32_t *refs;       // 2-bit type | 30-bit arena index
	const float4 *spheres;      // {centre, radius}
	const int *sphereMat;
	const float4 *boxes;        // 2 per box: {lo.xyz, hi.x}, {hi.y, hi.z, mat bits, 0}
	const float4 *affS;         // 3 per affine sphere: inverse rows with centre in .w
	const float4 *affSN;        // 3 per affine sphere: normal matrix rows, material bits
	const float4 *affB;
	const float4 *affBN;
	const float *planes;        // wire format, 26 floats per plane
	const float4 *materials;    // 4 float4 per material: diffuse, reflective, refractive, emittance+ior
};

// One scene's slices of the arena and its traversal parameters.
struct SceneRange {
	int rootCount;
	int roots[4];
	float rootBounds[4][6];
	int nodeBase;               // first arena slot of the scene's nodes
	int topSlots;               // slots from nodeBase holding exactly its top levels (breadth first)
	int refBase, refCount;
	int sphBase, sphCount, boxBase, boxCount, asBase, asCount, abBase, abCount;
	int planeBase, planeCount;
	int useBvh;                 // BVH traversal (1) or linear scan (0)
	int stackLimit;             // BVH_STACK_SIZE semantics
	int fastSlabs;              // every node and box bound is 0 or of magnitude in [2^-35, 2^58]
	float bg[3];
};

__device__ __forceinline__ MaterialF loadMaterial(const DevScene &S, int index) {
	MaterialF m;
	if (index < 0) {
		m.diffuse = m.reflective = m.refractive = m.emittance = mk(0.0f, 0.0f, 0.0f);
		m.ior = 0.0f;
		return m;
	}
	const float4 *p = S.materials + 4 * index;
	float4 a = p[0], b = p[1], c = p[2], d = p[3];
	m.diffuse = mk(a.x, a.y, a.z);
	m.reflective = mk(b.x, b.y, b.z);
	m.refractive = mk(c.x, c.y, c.z);
	m.emittance = mk(d.x, d.y, d.z);
	m.ior = d.w;
	return m;
}

// ------------------------------------------------------------ ray constants
//
// Every slab and sphere test of a closest-hit query divides by the same few
// values: a direction component, or dot(d, d).  The quotients must be the
// correctly rounded ones the reference computes, so instead of one IEEE
// division per slab (about eleven instructions with a quarter-rate reciprocal)
// a query computes the correctly rounded reciprocal y = RN(1/b) of each
// divisor once, and each quotient a/b as
//     q = RN(a y),  r = a - b q (exact, one fma),  RN(q + r y) (one fma),
// which equals RN(a/b) when y = RN(1/b) and nothing under- or overflows
// (P. Markstein, "Computation of elementary functions on the IBM RISC
// System/6000 processor", IBM J. Res. Dev. 34(1), 1990, Theorem 1).  The
// range kept is |a|, |b| in [2^-60, 2^60], where q, r and the result stay
// normal; a = 0 returns q = RN(0 y), which is 0/b with its sign.
//
// A slab numerator is a bound minus an origin coordinate.  When both are 0 or
// of magnitude in [2^-35, 2^58], their difference is 0 or a multiple of 2^-58
// below 2^59, so it lies in range without a test: the scene records whether
// its node and box bounds qualify, the ray whether its origin and direction
// do.  Rays that do not qualify (and a sphere or affine numerator, which is
// computed, when out of range) take the IEEE division.
struct Ray {
	f3 o, d;
	f3 inv;        // RN(1 / d.c), 0 where d.c == 0
	float dd;      // dot(d, d)
	float invDD;   // RN(1 / dd)
	bool fast;     // slab quotients of this ray may use the reciprocals
};
__device__ __forceinline__ bool magnitudeIn(float v, float lo, float hi) {
	const float a = fabsf(v);
	return a >= lo && a <= hi;
}
__device__ __forceinline__ Ray makeRay(f3 o, f3 d, bool sceneFast) {
	Ray r;
	r.o = o;
	r.d = d;
	r.inv = mk(d.x != 0.0f ? 1.0f / d.x : 0.0f, d.y != 0.0f ? 1.0f / d.y : 0.0f, d.z != 0.0f ? 1.0f / d.z : 0.0f);
	r.dd = dot(d, d);
	r.invDD = 1.0f / r.dd;
	auto dirOk = [](float v) { return v == 0.0f || magnitudeIn(v, 0x1p-60f, 0x1p60f); };
	auto origOk = [](float v) { return v == 0.0f || magnitudeIn(v, 0x1p-35f, 0x1p58f); };
	r.fast = sceneFast && dirOk(d.x) && dirOk(d.y) && dirOk(d.z) && magnitudeIn(r.dd, 0x1p-60f, 0x1p60f) &&
	         origOk(o.x) && origOk(o.y) && origOk(o.z);
	return r;
}
// Slab-quotient form of a call site: kSlabAny reads the ray's flag, the
// others are known at compile time (a traversal loop specialised per form).
constexpr int kSlabAny = 0, kSlabFast = 1, kSlabIeee = 2, kSlabFastNz = 3;
template <int kSlab>
__device__ __forceinline__ bool rayFast(const Ray &r) {
	if constexpr (kSlab == kSlabAny) return r.fast;
	else return kSlab != kSlabIeee;
}
// RN(num / den) from inv = RN(1 / den); num == 0 or |num| in [2^-60, 2^60].
__device__ __forceinline__ float exactQuotient(float num, float den, float inv) {
	const float q = num * inv;
	const float c = fmaf(fmaf(-q, den, num), inv, q);
	return num == 0.0f ? q : c;
}
__device__ __forceinline__ bool quotientOk(float num) { return num == 0.0f || magnitudeIn(num, 0x1p-60f, 0x1p60f); }
// num / den of a computed numerator: the reciprocal form when the ray allows
// it and num is in range, the IEEE division otherwise.
__device__ __forceinline__ float quotient(float num, float den, float inv, bool fast) {
	if (fast && quotientOk(num)) return exactQuotient(num, den, inv);
	return num / den;
}

// ------------------------------------------------------------ intersection
// Each test returns the reference's "found" predicate against `limit`
// (closestT in the reference) and writes the candidate distance.

// traceSphere body (GpuKernelSources.java:343-381).
struct SphereSolve { float t, radialOffset; f3 closest; };
// kFast: quotients by dd's reciprocal; `ok` is cleared if a numerator is out
// of range (the result is then discarded).
template <bool kFast>
__device__ __forceinline__ bool sphereSolveT(float4 s, const Ray &ray, SphereSolve &r, bool &ok) {
	auto div = [&](float num) {
		if constexpr (kFast) {
			ok = ok && quotientOk(num);
			return exactQuotient(num, ray.dd, ray.invDD);
		} else {
			return num / ray.dd;
		}
	};
	const f3 d = ray.d;
	f3 oc = ray.o - mk(s.x, s.y, s.z);
	float b = dot(oc, d);
	float centerT = div(-b);
	f3 cl = mk(fmaf(d.x, centerT, oc.x), fmaf(d.y, centerT, oc.y), fmaf(d.z, centerT, oc.z));
	float corr = div(-dot(cl, d));
	centerT += corr;
	cl = mk(fmaf(d.x, corr, cl.x), fmaf(d.y, corr, cl.y), fmaf(d.z, corr, cl.z));
	float rs = fmaf(s.w, s.w, -dot(cl, cl));
	if (rs <= 0.0f) return false;
	float halfSpan = sqrtf(div(rs));
	float t1 = centerT - halfSpan, t2 = centerT + halfSpan;
	if (t1 > kEps) { r.t = t1; r.radialOffset = -halfSpan; }
	else if (t2 > kEps) { r.t = t2; r.radialOffset = halfSpan; }
	else return false;
	r.closest = cl;
	return true;
}
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool sphereSolve(float4 s, const Ray &ray, SphereSolve &r) {
	bool ok = true;
	if (rayFast<kSlab>(ray)) {
		SphereSolve f;
		const bool hit = sphereSolveT<true>(s, ray, f, ok);
		if (ok) {
			r = f;
			return hit;
		}
	}
	return sphereSolveT<false>(s, ray, r, ok);
}
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool testSphere(float4 s, const Ray &ray, float limit, float &t) {
	SphereSolve r;
	if (!sphereSolve<kSlab>(s, ray, r)) return false;
	t = r.t;
	return r.t < limit;
}

// Slab of one axis; false when the zero-direction branch rejects
// (GpuKernelSources.java:476-515, :599-639).
__device__ __forceinline__ bool slab(float oc, float dc, float lo, float hi, float &tmin, float &tmax) {
	if (dc == 0.0f) {
		if (oc < lo || oc > hi) return false;
		tmin = -INFINITY; tmax = INFINITY;
		return true;
	}
	float t1 = (lo - oc) / dc, t2 = (hi - oc) / dc;
	tmin = fminf(t1, t2); tmax = fmaxf(t1, t2);
	return true;
}

// The same slab of a world-space box by the ray's reciprocal of dc.
__device__ __forceinline__ bool slabFast(float oc, float dc, float inv, float lo, float hi, float &tmin, float &tmax) {
	if (dc == 0.0f) {
		if (oc < lo || oc > hi) return false;
		tmin = -INFINITY; tmax = INFINITY;
		return true;
	}
	float t1 = exactQuotient(lo - oc, dc, inv), t2 = exactQuotient(hi - oc, dc, inv);
	tmin = fminf(t1, t2); tmax = fmaxf(t1, t2);
	return true;
}

// Enter/exit selection (GpuKernelSources.java:517-556, :641-680).
struct BoxSolve { float t; int axis; float sign; };
__device__ __forceinline__ bool boxSelect(float x0, float x1, float y0, float y1, float z0, float z1, float after, f3 d,
                                          BoxSolve &r) {
	float tEnter = fmaxf(fmaxf(x0, y0), z0);
	float tExit = fminf(fminf(x1, y1), z1);
	if (tEnter >= tExit || tExit <= after) return false;
	if (tEnter > after) {
		int axis = 0; float ev = x0;
		if (y0 > ev) { ev = y0; axis = 1; }
		if (z0 > ev) axis = 2;
		r.t = tEnter; r.axis = axis;
		r.sign = comp(d, axis) > 0.0f ? -1.0f : 1.0f;
	} else {
		int axis = 0; float xv = x1;
		if (y1 < xv) { xv = y1; axis = 1; }
		if (z1 < xv) axis = 2;
		r.t = tExit; r.axis = axis;
		r.sign = comp(d, axis) >= 0.0f ? 1.0f : -1.0f;
	}
	return true;
}
__device__ __forceinline__ f3 axisNormal(int axis, float sign) {
	return axis == 0 ? mk(sign, 0.0f, 0.0f) : (axis == 1 ? mk(0.0f, sign, 0.0f) : mk(0.0f, 0.0f, sign));
}

// traceBox body (GpuKernelSources.java:594-684).
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool boxSolve(float4 b0, float4 b1, const Ray &ray, BoxSolve &r) {
	float x0, x1, y0, y1, z0, z1;
	const f3 o = ray.o, d = ray.d;
	if (rayFast<kSlab>(ray)) {
		if (!slabFast(o.x, d.x, ray.inv.x, b0.x, b0.w, x0, x1)) return false;
		if (!slabFast(o.y, d.y, ray.inv.y, b0.y, b1.x, y0, y1)) return false;
		if (!slabFast(o.z, d.z, ray.inv.z, b0.z, b1.y, z0, z1)) return false;
	} else {
		if (!slab(o.x, d.x, b0.x, b0.w, x0, x1)) return false;
		if (!slab(o.y, d.y, b0.y, b1.x, y0, y1)) return false;
		if (!slab(o.z, d.z, b0.z, b1.y, z0, z1)) return false;
	}
	return boxSelect(x0, x1, y0, y1, z0, z1, kEps, d, r);
}
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool testBox(float4 b0, float4 b1, const Ray &ray, float limit, float &t) {
	BoxSolve r;
	if (!boxSolve<kSlab>(b0, b1, ray, r)) return false;
	t = r.t;
	return !(r.t >= limit);
}

// affineRayShift / applyAffine* (GpuKernelSources.java:276-301); rows hold
// the inverse linear map with the world centre in .w.
struct AffineLocal { float shift; f3 lo, ld; };
template <int kSlab = kSlabAny>
__device__ __forceinline__ void affineLocal(float4 r0, float4 r1, float4 r2, const Ray &ray, AffineLocal &a) {
	const f3 o = ray.o, d = ray.d;
	f3 c = mk(r0.w, r1.w, r2.w);
	float den = ray.dd;
	a.shift = den > 0.0f ? fmaxf(0.0f, quotient(dot(c - o, d), den, ray.invDD, rayFast<kSlab>(ray))) : 0.0f;
	f3 rel = mk(fmaf(d.x, a.shift, o.x - r0.w), fmaf(d.y, a.shift, o.y - r1.w), fmaf(d.z, a.shift, o.z - r2.w));
	a.lo = mk(r0.x * rel.x + r0.y * rel.y + r0.z * rel.z,
	          r1.x * rel.x + r1.y * rel.y + r1.z * rel.z,
	          r2.x * rel.x + r2.y * rel.y + r2.z * rel.z);
	a.ld = mk(r0.x * d.x + r0.y * d.y + r0.z * d.z,
	          r1.x * d.x + r1.y * d.y + r1.z * d.z,
	          r2.x * d.x + r2.y * d.y + r2.z * d.z);
}

// normalizedDirection (GpuKernelSources.java:317-336).
__device__ inline bool normalizedDirection(f3 v, f3 &unit, float &invLen) {
	float l2 = dot(v, v);
	if (isfinite(l2) && fabsf(l2 - 1.0f) <= 2.0e-6f) { unit = v; invLen = 1.0f; return true; }
	float scale = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
	if (!(scale > 0.0f) || !isfinite(scale)) return false;
	f3 s = mk(v.x / scale, v.y / scale, v.z / scale);
	float sl = sqrtf(dot(s, s));
	float dl = scale * sl;
	float inv = 1.0f / dl;
	if (!(inv > 0.0f) || !isfinite(inv)) return false;
	unit = s * (1.0f / sl);
	invLen = inv;
	return true;
}

// traceAffineSphere body (GpuKernelSources.java:408-443); returns the local
// parameter hitT and the world distance.
__device__ __forceinline__ bool affineSphereSolve(const AffineLocal &a, float &hitT, float &worldT) {
	float afterLocal = kEps - a.shift;
	float solveA = dot(a.ld, a.ld);
	f3 solveDir = a.ld;
	float pscale = 1.0f;
	if (!(solveA > 0.0f) || !isfinite(solveA)) {
		f3 u; float inv;
		if (!normalizedDirection(a.ld, u, inv)) return false;
		solveDir = u; solveA = 1.0f; pscale = inv;
	}
	float scaledAfterT = afterLocal / pscale;
	float centerT = -dot(a.lo, solveDir) / solveA;
	f3 cl = a.lo + solveDir * centerT;
	float rs = fmaf(-cl.x, cl.x, fmaf(-cl.y, cl.y, fmaf(-cl.z, cl.z, 1.0f)));
	if (rs <= 0.0f) return false;
	float halfSpan = sqrtf(rs / solveA);
	float t1 = centerT - halfSpan, t2 = centerT + halfSpan;
	float sh;
	if (t1 > scaledAfterT) sh = t1;
	else if (t2 > scaledAfterT) sh = t2;
	else return false;
	hitT = sh * pscale;
	worldT = a.shift + hitT;
	return true;
}
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool testAffineSphere(const float4 *rec, const Ray &ray, float limit, float &t) {
	AffineLocal a;
	affineLocal<kSlab>(rec[0], rec[1], rec[2], ray, a);
	float hitT, worldT;
	if (!affineSphereSolve(a, hitT, worldT)) return false;
	t = worldT;
	return !(worldT >= limit);
}

// traceAffineBox body (GpuKernelSources.java:470-558).
__device__ __forceinline__ bool affineBoxSolve(const AffineLocal &a, BoxSolve &r, float &worldT) {
	float x0, x1, y0, y1, z0, z1;
	if (!slab(a.lo.x, a.ld.x, -1.0f, 1.0f, x0, x1)) return false;
	if (!slab(a.lo.y, a.ld.y, -1.0f, 1.0f, y0, y1)) return false;
	if (!slab(a.lo.z, a.ld.z, -1.0f, 1.0f, z0, z1)) return false;
	if (!boxSelect(x0, x1, y0, y1, z0, z1, kEps - a.shift, a.ld, r)) return false;
	worldT = a.shift + r.t;
	return true;
}
template <int kSlab = kSlabAny>
__device__ __forceinline__ bool testAffineBox(const float4 *rec, const Ray &ray, float limit, float &t) {
	AffineLocal a;
	affineLocal<kSlab>(rec[0], rec[1], rec[2], ray, a);
	BoxSolve r;
	float worldT;
	if (!affineBoxSolve(a, r, worldT)) return false;
	t = worldT;
	return !(worldT >= limit);
}

// The cull bound of a query whose best distance is t: a box whose entry lies
// beyond it is skipped (GpuKernelSources.java:747, :878-880).  It changes only
// when a leaf improves the best hit, so traversal loops keep it in a register.
__device__ __forceinline__ float cullBound(float t) { return isfinite(t) ? fmaf(t, 1.0000015f, 1.0e-6f) : t; }

// bvhEntryDistance (GpuKernelSources.java:736-770) on 6 bounds, IEEE division;
// cb = cullBound(best distance).
__device__ __forceinline__ float entryDistanceIeee(float lx, float ly, float lz, float hx, float hy, float hz, f3 o,
                                                   f3 d, float cb) {
	float tMin = kEps;
	float tMax = cb;
#define MRAD_ENTRY_AXIS(OC, DC, LO, HI)                        \
	if (DC == 0.0f) {                                          \
		if (OC < LO || OC > HI) return INFINITY;               \
	} else {                                                   \
		float t0 = (LO - OC) / DC, t1 = (HI - OC) / DC;        \
		if (t0 > t1) { float tmp = t0; t0 = t1; t1 = tmp; }    \
		if (t1 > 0.0f) t1 *= 1.00000072f;                      \
		tMin = fmaxf(tMin, t0);                                \
		tMax = fminf(tMax, t1);                                \
		if (tMax < tMin) return INFINITY;                      \
	}
	MRAD_ENTRY_AXIS(o.x, d.x, lx, hx)
	MRAD_ENTRY_AXIS(o.y, d.y, ly, hy)
	MRAD_ENTRY_AXIS(o.z, d.z, lz, hz)
#undef MRAD_ENTRY_AXIS
	return tMin <= cb ? tMin : INFINITY;
}
// The same distance by the ray's reciprocals.  A zero quotient may come out
// with the other sign here (0 / b by the reciprocal form); no branch of the
// distance tells +0 from -0 (tMin >= 1e-4 > 0 and tMax > 0 before every axis),
// so the zero case needs no select.
__device__ __forceinline__ float entryDistanceFast(float lx, float ly, float lz, float hx, float hy, float hz,
                                                   const Ray &r, float cb) {
	float tMin = kEps;
	float tMax = cb;
#define MRAD_ENTRY_AXIS(OC, DC, INV, LO, HI)                                       \
	if (DC == 0.0f) {                                                              \
		if (OC < LO || OC > HI) return INFINITY;                                   \
	} else {                                                                       \
		const float n0 = LO - OC, n1 = HI - OC;                                    \
		const float q0 = n0 * INV, q1 = n1 * INV;                                  \
		float t0 = fmaf(fmaf(-q0, DC, n0), INV, q0);                               \
		float t1 = fmaf(fmaf(-q1, DC, n1), INV, q1);                               \
		if (t0 > t1) { const float tmp = t0; t0 = t1; t1 = tmp; }                  \
		if (t1 > 0.0f) t1 *= 1.00000072f;                                          \
		tMin = fmaxf(tMin, t0);                                                    \
		tMax = fminf(tMax, t1);                                                    \
		if (tMax < tMin) return INFINITY;                                          \
	}
	MRAD_ENTRY_AXIS(r.o.x, r.d.x, r.inv.x, lx, hx)
	MRAD_ENTRY_AXIS(r.o.y, r.d.y, r.inv.y, ly, hy)
	MRAD_ENTRY_AXIS(r.o.z, r.d.z, r.inv.z, lz, hz)
#undef MRAD_ENTRY_AXIS
	return tMin <= cb ? tMin : INFINITY;
}
// The reciprocal form for a ray with no zero direction component: the
// reference's zero-direction branch of each axis never runs.
__device__ __forceinline__ float entryDistanceFastNz(float lx, float ly, float lz, float hx, float hy, float hz,
                                                     const Ray &r, float cb) {
	float tMin = kEps;
	float tMax = cb;
#define MRAD_ENTRY_AXIS(OC, DC, INV, LO, HI)                                       \
	{                                                                              \
		const float n0 = LO - OC, n1 = HI - OC;                                    \
		const float q0 = n0 * INV, q1 = n1 * INV;                                  \
		float t0 = fmaf(fmaf(-q0, DC, n0), INV, q0);                               \
		float t1 = fmaf(fmaf(-q1, DC, n1), INV, q1);                               \
		if (t0 > t1) { const float tmp = t0; t0 = t1; t1 = tmp; }                  \
		if (t1 > 0.0f) t1 *= 1.00000072f;                                          \
		tMin = fmaxf(tMin, t0);                                                    \
		tMax = fminf(tMax, t1);                                                    \
		if (tMax < tMin) return INFINITY;                                          \
	}
	MRAD_ENTRY_AXIS(r.o.x, r.d.x, r.inv.x, lx, hx)
	MRAD_ENTRY_AXIS(r.o.y, r.d.y, r.inv.y, ly, hy)
	MRAD_ENTRY_AXIS(r.o.z, r.d.z, r.inv.z, lz, hz)
#undef MRAD_ENTRY_AXIS
	return tMin <= cb ? tMin : INFINITY;
}
// Branch-free forms of the two reciprocal entry distances: every axis is
// computed and the slab test is made once on the final interval.  The result
// is the same: the interval only shrinks, so it ends empty exactly when one
// of the per-axis checks would have returned; a zero direction component
// contributes no bounds, or an empty interval when the origin lies outside
// the slab (the reference's zero-direction branch).
//
// The far-slab widening f(t) = t > 0 ? RN(t * 1.00000072) : t is applied
// once, to the least far distance of the three axes, instead of per axis:
// f is non-decreasing (RN of a product with a positive constant is
// monotone, and f(t) > 0 >= f(t') for t > 0 >= t'), so
// min(cb, f(b.x), f(b.y), f(b.z)) = min(cb, f(min(b.x, b.y, b.z))).
#ifndef MRAD_ENTRY_LEAN
#define MRAD_ENTRY_LEAN 1
#endif
// The near (a) and far (b) slab distances of one axis.
__device__ __forceinline__ void slabAxisBf(float oc, float dc, float inv, float lo, float hi, bool nz, float &a,
                                           float &b) {
	const float n0 = lo - oc, n1 = hi - oc;
	const float q0 = n0 * inv, q1 = n1 * inv;
	const float t0 = fmaf(fmaf(-q0, dc, n0), inv, q0);
	const float t1 = fmaf(fmaf(-q1, dc, n1), inv, q1);
	a = fminf(t0, t1);
	b = fmaxf(t0, t1);
	if (!nz) {
		const bool zero = dc == 0.0f;
		const bool outside = oc < lo || oc > hi;
		a = zero ? (outside ? INFINITY : -INFINITY) : a;
		b = zero ? INFINITY : b;
	}
}
template <bool kNz>
__device__ __forceinline__ float entryDistanceFastBf(float lx, float ly, float lz, float hx, float hy, float hz,
                                                     const Ray &r, float cb) {
	float ax, bx, ay, by, az, bz;
	slabAxisBf(r.o.x, r.d.x, r.inv.x, lx, hx, kNz, ax, bx);
	slabAxisBf(r.o.y, r.d.y, r.inv.y, ly, hy, kNz, ay, by);
	slabAxisBf(r.o.z, r.d.z, r.inv.z, lz, hz, kNz, az, bz);
	// kEps enters through a two-operand max (an inline literal): a
	// three-operand max would hold it in a scalar register, which this
	// kernel's scalar budget pays for with spill reloads inside the loop
	const float tMin = fmaxf(fmaxf(fmaxf(kEps, ax), ay), az);
	const float bMin = fminf(fminf(bx, by), bz);
#if MRAD_ENTRY_LEAN
	// Two checks of the reference's form are dropped because they cannot
	// change the result here (no value below is NaN: the bounds, origin and
	// reciprocals are finite, and a zero direction component yields +-inf
	// through selects only):
	// * the widening's sign test: for bMin <= 0 the unwidened and the widened
	//   far distance are both <= 0 < kEps <= tMin, so the box is culled
	//   either way; for bMin > 0 the product is the reference's;
	// * the final tMin <= cb: tMax = min(cb, ..) >= tMin already implies it.
	const float tMax = fminf(cb, bMin * 1.00000072f);
	return tMax < tMin ? INFINITY : tMin;
#else
	const float tMax = fminf(cb, bMin > 0.0f ? bMin * 1.00000072f : bMin);
	return (tMax < tMin || !(tMin <= cb)) ? INFINITY : tMin;
#endif
}
#ifndef MRAD_ENTRY_BRANCH_FREE
#define MRAD_ENTRY_BRANCH_FREE 1
#endif
// bvhEntryDistance (GpuKernelSources.java:736-770) on 6 bounds with the cull
// bound cb = cullBound(best distance).  kSlab picks the quotient form:
// kSlabAny by the ray's flag, kSlabFast or kSlabIeee when the caller already
// knows it (a traversal loop specialised per form keeps one form on its hot
// path).
template <int kSlab = kSlabAny>
__device__ __forceinline__ float entryDistanceCb(float lx, float ly, float lz, float hx, float hy, float hz,
                                                 const Ray &r, float cb) {
#if MRAD_ENTRY_BRANCH_FREE
	if constexpr (kSlab == kSlabFastNz) return entryDistanceFastBf<true>(lx, ly, lz, hx, hy, hz, r, cb);
	if constexpr (kSlab == kSlabFast) return entryDistanceFastBf<false>(lx, ly, lz, hx, hy, hz, r, cb);
#endif
	if constexpr (kSlab == kSlabFastNz) return entryDistanceFastNz(lx, ly, lz, hx, hy, hz, r, cb);
	if constexpr (kSlab == kSlabFast) return entryDistanceFast(lx, ly, lz, hx, hy, hz, r, cb);
	if constexpr (kSlab == kSlabIeee) return entryDistanceIeee(lx, ly, lz, hx, hy, hz, r.o, r.d, cb);
	if (r.fast) return entryDistanceFast(lx, ly, lz, hx, hy, hz, r, cb);
	return entryDistanceIeee(lx, ly, lz, hx, hy, hz, r.o, r.d, cb);
}
// The same against the best distance itself.
template <int kSlab = kSlabAny>
__device__ __forceinline__ float entryDistance(float lx, float ly, float lz, float hx, float hy, float hz, const Ray &r,
                                               float beforeT) {
	return entryDistanceCb<kSlab>(lx, ly, lz, hx, hy, hz, r, cullBound(beforeT));
}

// ------------------------------------------------------------ closest hit

// Best candidate: distance, rank (type << 30 | arena index; planes use kind 4
// and never enter a rank comparison), kind and arena index.
struct Best { float t; uint32_t rank; int kind; int index; };

template <bool kMetrics>
struct Counters {
	uint32_t v[kStatCount];
	__device__ __forceinline__ void zero() {
#pragma unroll
		for (int i = 0; i < kStatCount; i++) v[i] = 0;
	}
	__device__ __forceinline__ void add(int slot, uint32_t n) { v[slot] += n; }
	__device__ __forceinline__ void maxStack(int size) { v[9] = v[9] > (uint32_t) size ? v[9] : (uint32_t) size; }
};
template <>
struct Counters<false> {
	__device__ __forceinline__ void zero() {}
	__device__ __forceinline__ void add(int, uint32_t) {}
	__device__ __forceinline__ void maxStack(int) {}
};

__device__ __forceinline__ int typeSlot(uint32_t type) {
	return type == kRefBox ? 1 : type == kRefAffineSphere ? 3 : type == kRefAffineBox ? 4 : 0;
}

// tracePrimitiveRef with the reference's (t, order) acceptance
// (GpuKernelSources.java:772-806) on the primitive's test record `rec`.
template <int kSlab = kSlabAny>
__device__ __forceinline__ void testRefAt(const float4 *rec, uint32_t ref, uint32_t type, const Ray &ray, Best &best) {
	const int index = (int) (ref & kRefIndexMask);
	const float limit = nextafterf(best.t, INFINITY);
	float t = 0.0f;
	bool found;
	int kind;
	if (type == kRefBox) {
		kind = kKindBox;
		found = testBox<kSlab>(rec[0], rec[1], ray, limit, t);
	} else if (type == kRefAffineSphere) {
		kind = kKindAffineSphere;
		found = testAffineSphere<kSlab>(rec, ray, limit, t);
	} else if (type == kRefAffineBox) {
		kind = kKindAffineBox;
		found = testAffineBox<kSlab>(rec, ray, limit, t);
	} else {
		kind = kKindSphere;
		found = testSphere<kSlab>(rec[0], ray, limit, t);
	}
	const uint32_t rank = type | (uint32_t) index;
	if (!found || t > best.t || (t == best.t && rank >= best.rank)) return;
	best.t = t; best.rank = rank; best.kind = kind; best.index = index;
}
__device__ __forceinline__ const float4 *primitiveRecord(const DevScene &S, uint32_t ref, uint32_t type) {
	const size_t index = ref & kRefIndexMask;
	return type == kRefBox ? S.boxes + 2 * index : type == kRefAffineSphere ? S.affS + 3 * index
	     : type == kRefAffineBox ? S.affB + 3 * index : S.spheres + index;
}
__device__ __forceinline__ void testRef(const DevScene &S, uint32_t ref, uint32_t type, const Ray &ray, Best &best) {
	testRefAt(primitiveRecord(S, ref, type), ref, type, ray, best);
}

// The refs of a leaf node record `rec` with data `nd` = {kind, ref if one,
// first ref, count}.  A one-ref leaf carries its primitive's test record in
// the node record (rec[0..2]), so only the node itself is fetched.
template <bool kMetrics, int kSlab = kSlabAny>
__device__ __forceinline__ void testLeaf(const DevScene &S, const SceneRange &R, const float4 *rec, int4 nd,
                                         const Ray &ray, Best &best, Counters<kMetrics> &c) {
	c.add(11, 1);
	c.add(nd.x == -1 ? 13 : 12, 1);
	c.add(2, (uint32_t) nd.w);
	const int kind = nd.x;
	const bool mixed = kind > -2 || kind < -5;
	const uint32_t leafType = kind == -2 ? kRefSphere : kind == -3 ? kRefBox
	                        : kind == -4 ? kRefAffineSphere : kRefAffineBox;
	const bool inlined = nd.w == 1;
	for (int i = 0; i < nd.w; i++) {
		const int ri = nd.z + i;
		if (!inlined && (ri < R.refBase || ri >= R.refBase + R.refCount)) continue;
		const uint32_t ref = inlined ? (uint32_t) nd.y : S.refs[ri];
		const uint32_t type = mixed ? (ref & kRefTypeMask) : leafType;
		c.add(3 + typeSlot(type), 1);
		testRefAt<kSlab>(inlined ? rec : primitiveRecord(S, ref, type), ref, type, ray, best);
	}
}

// Traversal stack storage.  The top kShortStack entries of each lane live in
// LDS, lane-interleaved (entry i of lane l at [i * 64 + l]) so that a wave's
// accesses hit 64 distinct banks whatever each lane's depth; deeper entries
// spill to a thread-private array.  Capacity and overflow semantics are the
// reference's BVH_STACK_SIZE (GpuKernelSources.java:858-859, :871, :924, :933).
// Leaf batching of the traversal loop (traceBvh's kBatch): 0 tests a leaf's
// refs as soon as a lane pops it; K > 0 holds popped leaves until K lanes of
// the wave (or all that still traverse) hold one.  Both visit nodes and test
// refs in the same per-lane order, so hits and counters are identical.
// 1 specialises the traversal loop per slab-quotient form (see traceClosest);
// 2 adds a loop for reciprocal-form rays with no zero direction component.
#ifndef MRAD_SPLIT_SLAB
#define MRAD_SPLIT_SLAB 2
#endif
// 1 carries the nearer child to the next iteration in registers instead of
// pushing and popping it.  With the packed stack, the branch-free push and
// the select-ordered children, pushing both is faster in every kernel
// (megakernel 1.6 %, adjoint backward 2.3 %, wavefront trace 3.1 %): the
// carried and the popping lanes of a wave take different paths.
#ifndef MRAD_CARRY_NEAR
#define MRAD_CARRY_NEAR 0
#endif
// 1 fetches a popped node's whole 64-byte record before its kind is known.
#ifndef MRAD_NODE_PREFETCH
#define MRAD_NODE_PREFETCH 0
#endif
// 1 pushes an interior node's children without a branch per child: both
// slots are written and the stack size advances by the finite ones.
#ifndef MRAD_PUSH_BRANCH_FREE
#define MRAD_PUSH_BRANCH_FREE 1
#endif
#ifndef MRAD_SHORT_STACK
#define MRAD_SHORT_STACK 10
#endif
constexpr int kShortStack = MRAD_SHORT_STACK;

// The pointers are typed as LDS (address space 3).  With generic pointers
// the compiler merges the LDS and spill reads of TravStack into one flat load
// through a select, which is slower than a ds_read and waits for every
// outstanding global load (megakernel 1.7 % slower on the same GPU).
#ifndef MRAD_LDS_STACK_TYPED
#define MRAD_LDS_STACK_TYPED 1
#endif
#if MRAD_LDS_STACK_TYPED
typedef __attribute__((address_space(3))) int LdsInt;
typedef __attribute__((address_space(3))) float LdsFloat;
typedef __attribute__((address_space(3))) unsigned long long LdsSlot;
#else
typedef int LdsInt;
typedef float LdsFloat;
typedef unsigned long long LdsSlot;
#endif
// 1 keeps a stack entry's node and entry distance in one 64-bit slot, so a
// pop is one LDS (or spill) read behind one LDS-or-spill branch instead of
// two of each.
#ifndef MRAD_STACK_PACKED
#define MRAD_STACK_PACKED 1
#endif
struct LdsStack {
#if MRAD_STACK_PACKED
	LdsSlot *slot;    // lane base, stride 64: node | entry bits << 32
#else
	LdsInt *node;     // lane base, stride 64
	LdsFloat *entry;
#endif
};

// kShort is the number of top entries held in LDS (the LDS split); kCap the
// total storage capacity.
template <int kCap, int kShort = kShortStack>
struct TravStack {
	// kCap + 1 slots: a branch-free push may write one slot past the limit
	static constexpr int kSpill = kCap + 1 > kShort ? kCap + 1 - kShort : 1;
	LdsStack lds;
#if MRAD_STACK_PACKED
	unsigned long long spill[kSpill];
	__device__ __forceinline__ void set(int i, int n, float e) {
		const unsigned long long v = (uint32_t) n | (unsigned long long) __float_as_uint(e) << 32;
		if (kShort > 0 && i < kShort) lds.slot[i * 64] = v;
		else spill[i - kShort] = v;
	}
	__device__ __forceinline__ void get(int i, int &n, float &e) const {
		const unsigned long long v = (kShort > 0 && i < kShort) ? lds.slot[i * 64] : spill[i - kShort];
		n = (int) (uint32_t) v;
		e = __uint_as_float((uint32_t) (v >> 32));
	}
#else
	int spillNode[kSpill];
	float spillEntry[kSpill];
	__device__ __forceinline__ void set(int i, int n, float e) {
		if (kShort > 0 && i < kShort) { lds.node[i * 64] = n; lds.entry[i * 64] = e; }
		else { spillNode[i - kShort] = n; spillEntry[i - kShort] = e; }
	}
	__device__ __forceinline__ void get(int i, int &n, float &e) const {
		n = (kShort > 0 && i < kShort) ? lds.node[i * 64] : spillNode[i - kShort];
		e = (kShort > 0 && i < kShort) ? lds.entry[i * 64] : spillEntry[i - kShort];
	}
#endif
	__device__ __forceinline__ int node(int i) const { int n; float e; get(i, n, e); return n; }
	__device__ __forceinline__ float entry(int i) const { int n; float e; get(i, n, e); return e; }
};

// An LDS copy of the first `count` node slots of a scene (from arena slot
// `base`): the top levels of its BVH, laid out breadth first on upload.  A
// traversal with kTop > 0 reads those nodes from here instead of HBM.
typedef float Vec4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) Vec4f LdsFloat4;
struct TopNodes {
	const LdsFloat4 *p;
	int base, count;
};
__device__ __forceinline__ float4 ldsLoad4(const LdsFloat4 *q) {
	const Vec4f v = *q;
	return make_float4(v.x, v.y, v.z, v.w);
}

// traceBvh (GpuKernelSources.java:844-961) in three parts: bvhBegin inserts
// the roots, traceBvh's loop pops and expands stack entries, bvhEnd runs the
// overflow fallback.  Same stack discipline, same counters.  kCap is the
// storage capacity; the runtime limit R.stackLimit (<= kCap) gives
// BVH_STACK_SIZE semantics.
template <int kCap, int kShort = kShortStack>
struct BvhState {
	TravStack<kCap, kShort> st;
	int size;
	bool overflow;
};

template <bool kMetrics, int kCap, int kShort, int kSlab = kSlabAny>
__device__ __forceinline__ void bvhBegin(const SceneRange &R, const Ray &ray, const Best &best, Counters<kMetrics> &c,
                                         BvhState<kCap, kShort> &b) {
	b.size = 0;
	b.overflow = false;
	if (R.refCount <= 0 || R.rootCount <= 0 || R.rootCount > 4) return;
	const int cap = R.stackLimit;
	for (int r = 0; r < R.rootCount; r++) {
		const int ri = R.roots[r];
		if (ri < 0) { b.overflow = true; break; }
		c.add(1, 1);
		const float *rb = R.rootBounds[r];
		float e = entryDistance<kSlab>(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], ray, best.t);
		if (!isfinite(e)) continue;
		if (b.size >= cap) { b.overflow = true; break; }
		int pos = b.size;
		while (pos > 0 && e >= b.st.entry(pos - 1)) {
			b.st.set(pos, b.st.node(pos - 1), b.st.entry(pos - 1));
			pos--;
		}
		b.st.set(pos, ri, e);
		b.size++;
		c.maxStack(b.size);
	}
}

// The reference's stack-overflow fallback: every ref of the scene, linearly.
template <bool kMetrics, int kCap, int kShort>
__device__ __forceinline__ void bvhEnd(const DevScene &S, const SceneRange &R, const Ray &ray, Best &best,
                                       Counters<kMetrics> &c, const BvhState<kCap, kShort> &b) {
	if (!b.overflow) return;
	c.add(8, 1);
	for (int i = R.refBase; i < R.refBase + R.refCount; i++) {
		const uint32_t ref = S.refs[i];
		c.add(2, 1);
		c.add(3 + typeSlot(ref & kRefTypeMask), 1);
		testRef(S, ref, ref & kRefTypeMask, ray, best);
	}
}

// The children of an interior node in visiting order, the nearer first and
// the left one on a tie (GpuKernelSources.java:915-922), as selects rather
// than a swap under a branch.
__device__ __forceinline__ void orderChildren(int4 nd, float le, float re, int &nearN, int &farN, float &nearE,
                                              float &farE) {
	const bool swap = re < le;
	nearN = swap ? nd.y : nd.x;
	farN = swap ? nd.x : nd.y;
	nearE = swap ? re : le;
	farE = swap ? le : re;
}

// The default traversal loop of traceBvh (GpuKernelSources.java:895-958):
// pops and expands stack entries until the stack is empty or overflows, or,
// with kBudget, until `budget` pops have been made.  The query's state
// (stack, size, overflow, best) is then left in `b` and `best` as the next
// pop needs it, so a later call continues the same query with the same
// pops, tests and counters.
template <bool kMetrics, int kCap, int kShort, int kSlab, int kTop, bool kBudget>
__device__ __forceinline__ void bvhLoop(const DevScene &S, const SceneRange &R, const Ray &ray, Best &best,
                                        Counters<kMetrics> &c, BvhState<kCap, kShort> &b, TopNodes top,
                                        int budget) {
	const int cap = R.stackLimit;
	int size = b.size;
	bool overflow = b.overflow;
	float ch = cullBound(best.t);   // changes only when a leaf improves best.t
	while (!overflow && size > 0) {
		if constexpr (kBudget) {
			if (budget <= 0) break;
			budget--;
		}
		--size;
		int node;
		float entry;
		b.st.get(size, node, entry);
		if (entry > ch) continue;
		const float4 *rec = S.nodes + 4 * (size_t) (uint32_t) node;
#if MRAD_NODE_PREFETCH
		const float4 a = rec[0], bb = rec[1], cc = rec[2];
		const int4 nd = *reinterpret_cast<const int4 *>(rec + 3);
#else
		float4 a, bb, cc;
		int4 nd;
		const uint32_t rel = (uint32_t) node - (uint32_t) top.base;
		const bool inTop = kTop > 0 && rel < (uint32_t) top.count;
		if (inTop) {
			const LdsFloat4 *q = top.p + 4 * rel;
			const float4 w = ldsLoad4(q + 3);
			nd = make_int4(__float_as_int(w.x), __float_as_int(w.y), __float_as_int(w.z), __float_as_int(w.w));
			a = ldsLoad4(q);
			bb = ldsLoad4(q + 1);
			cc = ldsLoad4(q + 2);
		} else {
			nd = *reinterpret_cast<const int4 *>(rec + 3);
		}
#endif
		if (nd.x < 0) {
			testLeaf<kMetrics, kSlab>(S, R, rec, nd, ray, best, c);
			ch = cullBound(best.t);
			continue;
		}
		c.add(10, 1);
#if !MRAD_NODE_PREFETCH
		if (!inTop) {
			a = rec[0];
			bb = rec[1];
			cc = rec[2];
		}
#endif
		const float le = entryDistanceCb<kSlab>(a.x, a.y, a.z, a.w, bb.x, bb.y, ray, ch);
		const float re = entryDistanceCb<kSlab>(bb.z, bb.w, cc.x, cc.y, cc.z, cc.w, ray, ch);
		c.add(1, 2);
		int nearN, farN;
		float nearE, farE;
		orderChildren(nd, le, re, nearN, farN, nearE, farE);
#if MRAD_PUSH_BRANCH_FREE
		// Both children are written unconditionally, the nearer one above the
		// farther if that is finite; a non-finite child's slot lies past the
		// new top (storage holds cap + 1 entries) and is never read.  size <=
		// cap - 1 after the pop, so only two finite children overflow, after
		// the far push brought the maximum to cap.
		const int s1 = size + (isfinite(farE) ? 1 : 0);
		const int s2 = s1 + (isfinite(nearE) ? 1 : 0);
		if (s2 > cap) { c.maxStack(cap); overflow = true; break; }
		b.st.set(size, farN, farE);
		b.st.set(s1, nearN, nearE);
		size = s2;
		c.maxStack(size);
#else
		if (isfinite(farE)) {
			if (size >= cap) { overflow = true; break; }
			b.st.set(size, farN, farE);
			size++;
			c.maxStack(size);
		}
		if (isfinite(nearE)) {
			if (size >= cap) { overflow = true; break; }
			b.st.set(size, nearN, nearE);
			size++;
			c.maxStack(size);
		}
#endif
	}
	b.size = size;
	b.overflow = overflow;
}

template <bool kMetrics, int kCap, int kShort = kShortStack, int kSlab = kSlabAny, int kBatch = 0, int kTop = 0>
__device__ inline void traceBvh(const DevScene &S, const SceneRange &R, const Ray &ray, Best &best,
                                Counters<kMetrics> &c, LdsStack lds, TopNodes top = {}) {
	BvhState<kCap, kShort> b;
	b.st.lds = lds;
	bvhBegin<kMetrics, kCap, kShort, kSlab>(R, ray, best, c, b);
	const int cap = R.stackLimit;
	int size = b.size;
	bool overflow = b.overflow;
	float ch = cullBound(best.t);   // changes only when a leaf improves best.t
	if constexpr (kBatch > 0) {
	// The same per-lane sequence (pop, then expand an interior node or test a
	// leaf's refs before the next pop), but a lane that pops a leaf holds it
	// until kBatch lanes of the wave hold one (or every lane still traversing
	// does), and those leaves are tested together: the leaf branch then runs
	// with many active lanes instead of one or two.
	bool active = !overflow && size > 0;
	bool pending = false;
	int4 leaf = make_int4(0, 0, 0, 0);
	int leafNode = 0;
	while (active) {
		if (!pending) {
			--size;
			int node;
			float entry;
			b.st.get(size, node, entry);
			if (entry <= ch) {
				const float4 *rec = S.nodes + 4 * (size_t) (uint32_t) node;
				const int4 nd = *reinterpret_cast<const int4 *>(rec + 3);
				if (nd.x < 0) {
					leaf = nd;
					leafNode = node;
					pending = true;
				} else {
					c.add(10, 1);
					const float4 a = rec[0], bb = rec[1], cc = rec[2];
					const float le = entryDistanceCb<kSlab>(a.x, a.y, a.z, a.w, bb.x, bb.y, ray, ch);
					const float re = entryDistanceCb<kSlab>(bb.z, bb.w, cc.x, cc.y, cc.z, cc.w, ray, ch);
					c.add(1, 2);
					int nearN, farN;
					float nearE, farE;
					orderChildren(nd, le, re, nearN, farN, nearE, farE);
					// the default loop's branch-free push (same contents, maximum
					// and overflow point)
					const int s1 = size + (isfinite(farE) ? 1 : 0);
					const int s2 = s1 + (isfinite(nearE) ? 1 : 0);
					if (s2 > cap) {
						c.maxStack(cap);
						overflow = true;
					} else {
						b.st.set(size, farN, farE);
						b.st.set(s1, nearN, nearE);
						size = s2;
						c.maxStack(size);
					}
				}
			}
		}
		const unsigned long long waiting = __ballot(pending);
		if (waiting == __ballot(1) || __popcll(waiting) >= kBatch) {
			if (pending) {
				testLeaf<kMetrics, kSlab>(S, R, S.nodes + 4 * (size_t) leafNode, leaf, ray, best, c);
				ch = cullBound(best.t);
				pending = false;
			}
		}
		active = !overflow && (pending || size > 0);
	}
	} else {
#if MRAD_CARRY_NEAR
	// The nearer child is carried to the next iteration in registers instead
	// of being pushed and popped at once.  Same order and counters: the
	// overflow check and the stack-size maximum are taken as if it were
	// pushed, and its entry needs no cull check (it was computed against the
	// same best distance, so it does not exceed the cull bound).
	bool carry = false;
	int carryNode = 0;
	while (!overflow && (carry || size > 0)) {
		int node;
		if (carry) {
			node = carryNode;
			carry = false;
		} else {
			--size;
			float entry;
			b.st.get(size, node, entry);
			if (entry > ch) continue;
		}
		const float4 *rec = S.nodes + 4 * (size_t) (uint32_t) node;
		const int4 nd = *reinterpret_cast<const int4 *>(rec + 3);
		if (nd.x < 0) {
			testLeaf<kMetrics, kSlab>(S, R, rec, nd, ray, best, c);
			ch = cullBound(best.t);
			continue;
		}
		c.add(10, 1);
		const float4 a = rec[0], bb = rec[1], cc = rec[2];
		const float le = entryDistanceCb<kSlab>(a.x, a.y, a.z, a.w, bb.x, bb.y, ray, ch);
		const float re = entryDistanceCb<kSlab>(bb.z, bb.w, cc.x, cc.y, cc.z, cc.w, ray, ch);
		c.add(1, 2);
		int nearN, farN;
		float nearE, farE;
		orderChildren(nd, le, re, nearN, farN, nearE, farE);
#if MRAD_PUSH_BRANCH_FREE
		// size <= cap - 1 here (a pop, or a carry whose push left room), so
		// the far slot is in storage; a non-finite far child is written
		// past the top and never read.  Overflow exactly when the carried
		// child would not fit: the far push fits, so the maximum reaches cap.
		const int s1 = size + (isfinite(farE) ? 1 : 0);
		const bool pn = isfinite(nearE);
		if (pn && s1 >= cap) { c.maxStack(cap); overflow = true; break; }
		b.st.set(size, farN, farE);
		size = s1;
		c.maxStack(s1 + (pn ? 1 : 0));
		carry = pn;
		carryNode = nearN;
#else
		if (isfinite(farE)) {
			if (size >= cap) { overflow = true; break; }
			b.st.set(size, farN, farE);
			size++;
			c.maxStack(size);
		}
		if (isfinite(nearE)) {
			if (size >= cap) { overflow = true; break; }
			c.maxStack(size + 1);
			carry = true;
			carryNode = nearN;
		}
#endif
	}
#else
	b.size = size;
	b.overflow = overflow;
	bvhLoop<kMetrics, kCap, kShort, kSlab, kTop, false>(S, R, ray, best, c, b, top, 0);
	size = b.size;
	overflow = b.overflow;
#endif
	}
	b.size = size;
	b.overflow = overflow;
	bvhEnd<kMetrics>(S, R, ray, best, c, b);
}

// Linear scan of every primitive class in global order, strict against the
// running best (GpuKernelSources.java:993-1013).
template <bool kMetrics>
__device__ inline void traceLinear(const DevScene &S, const SceneRange &R, const Ray &ray, Best &best,
                                   Counters<kMetrics> &c) {
	c.add(2, (uint32_t) (R.sphCount + R.boxCount + R.asCount + R.abCount));
	c.add(3, (uint32_t) R.sphCount);
	c.add(4, (uint32_t) R.boxCount);
	c.add(6, (uint32_t) R.asCount);
	c.add(7, (uint32_t) R.abCount);
	float t;
	for (int i = R.sphBase; i < R.sphBase + R.sphCount; i++)
		if (testSphere(S.spheres[i], ray, best.t, t)) best = {t, kRefSphere | (uint32_t) i, kKindSphere, i};
	for (int i = R.boxBase; i < R.boxBase + R.boxCount; i++)
		if (testBox(S.boxes[2 * i], S.boxes[2 * i + 1], ray, best.t, t)) best = {t, kRefBox | (uint32_t) i, kKindBox, i};
	for (int i = R.asBase; i < R.asBase + R.asCount; i++)
		if (testAffineSphere(S.affS + 3 * i, ray, best.t, t))
			best = {t, kRefAffineSphere | (uint32_t) i, kKindAffineSphere, i};
	for (int i = R.abBase; i < R.abBase + R.abCount; i++)
		if (testAffineBox(S.affB + 3 * i, ray, best.t, t)) best = {t, kRefAffineBox | (uint32_t) i, kKindAffineBox, i};
}

// tracePlane (GpuKernelSources.java:708-734) distance part.
__device__ inline void tracePlanes(const DevScene &S, const SceneRange &R, f3 o, f3 d, Best &best) {
	for (int i = R.planeBase; i < R.planeBase + R.planeCount; i++) {
		const float *p = S.planes + 26 * i;
		f3 n = mk(p[0], p[1], p[2]);
		float denom = dot(n, d);
		if (denom == 0.0f) continue;
		float t = -(dot(n, o) + p[3]) / denom;
		if (t <= kEps || t >= best.t) continue;
		best = {t, 0xFFFFFFFFu, kKindPlane, i};
	}
}

// The reference's global primitive order of the best hit (GpuKernelSources.java:784-799, :1017).
__device__ __forceinline__ int localOrder(const SceneRange &R, const Best &b) {
	switch (b.kind) {
		case kKindSphere: return b.index - R.sphBase;
		case kKindBox: return R.sphCount + b.index - R.boxBase;
		case kKindAffineSphere: return R.sphCount + R.boxCount + b.index - R.asBase;
		case kKindAffineBox: return R.sphCount + R.boxCount + R.asCount + b.index - R.abBase;
		default: return R.sphCount + R.boxCount + R.asCount + R.abCount + b.index - R.planeBase;
	}
}

// choosePlaneMaterialIndex (GpuKernelSources.java:251-269).
__device__ inline int planeMaterialIndex(const float *p, f3 hp) {
	const float *g = p + 18;
	const int surface = (int) g[5];
	if (g[0] == 0.0f) return surface;
	f3 rel = hp - mk(p[4], p[5], p[6]);
	float b_e = dot(rel, mk(p[7], p[8], p[9])) / p[13];
	float b_f = dot(rel, mk(p[10], p[11], p[12])) / p[14];
	float u = (b_e - b_f * p[16]) / p[17];
	float v = (b_f - b_e * p[15]) / p[17];
	float hx = g[3] * 0.5f, hy = g[4] * 0.5f;
	float mx = (u + hx) - g[1] * floorf((u + hx) / g[1]);
	float my = (v + hy) - g[2] * floorf((v + hy) / g[2]);
	if (mx < g[3] || my < g[4]) return (int) g[7];
	return (int) g[6];
}

struct HitDetail { f3 pos, normal; int matIndex; };

// Normal, position and material of the winning primitive, recomputed with
// the arithmetic of the routine that accepted it.  Material indices in the
// arena are already rebased to the arena's material table.
__device__ inline HitDetail resolveHit(const DevScene &S, f3 o, f3 d, float t, int kind, int i) {
	HitDetail h;
	h.pos = o + d * t;
	const Ray ray = makeRay(o, d, false);
	if (kind == kKindSphere) {
		SphereSolve r;
		sphereSolve(S.spheres[i], ray, r);
		h.normal = normalizeSafe(mk(fmaf(d.x, r.radialOffset, r.closest.x), fmaf(d.y, r.radialOffset, r.closest.y),
		                            fmaf(d.z, r.radialOffset, r.closest.z)));
		h.matIndex = S.sphereMat[i];
	} else if (kind == kKindBox) {
		float4 b1 = S.boxes[2 * i + 1];
		BoxSolve r;
		boxSolve(S.boxes[2 * i], b1, ray, r);
		h.normal = axisNormal(r.axis, r.sign);
		h.matIndex = __float_as_int(b1.z);
	} else if (kind == kKindAffineSphere || kind == kKindAffineBox) {
		const bool isBox = kind == kKindAffineBox;
		const float4 *rec = (isBox ? S.affB : S.affS) + 3 * i;
		const float4 *nrm = (isBox ? S.affBN : S.affSN) + 3 * i;
		AffineLocal a;
		affineLocal(rec[0], rec[1], rec[2], ray, a);
		f3 local;
		if (isBox) {
			BoxSolve r;
			float worldT;
			affineBoxSolve(a, r, worldT);
			local = axisNormal(r.axis, r.sign);
		} else {
			float hitT, worldT;
			affineSphereSolve(a, hitT, worldT);
			local = a.lo + a.ld * hitT;
		}
		float4 n0 = nrm[0], n1 = nrm[1], n2 = nrm[2];
		f3 wn = mk(n0.x * local.x + n0.y * local.y + n0.z * local.z,
		           n0.w * local.x + n1.x * local.y + n1.y * local.z,
		           n1.z * local.x + n1.w * local.y + n2.x * local.z);
		h.normal = normalizeTransformedNormal(wn);
		h.matIndex = __float_as_int(n2.y);
	} else {
		const float *p = S.planes + 26 * i;
		h.normal = mk(p[0], p[1], p[2]);
		h.matIndex = planeMaterialIndex(p, h.pos);
	}
	return h;
}

// traceClosest (GpuKernelSources.java:963-1020).
template <bool kMetrics, int kCap, int kShort = kShortStack, int kBatch = 0, int kTop = 0>
__device__ __forceinline__ bool traceClosest(const DevScene &S, const SceneRange &R, f3 o, f3 d, Best &best,
                                             Counters<kMetrics> &c, LdsStack lds, TopNodes top = {}) {
	best.t = 1e30f;
	best.rank = 0xFFFFFFFFu;
	best.kind = -1;
	best.index = -1;
	const Ray ray = makeRay(o, d, R.fastSlabs != 0);
#if MRAD_SPLIT_SLAB
	// One traversal loop per slab-quotient form, so the common (reciprocal)
	// form's loop carries no per-test branch to the IEEE form.
	if (R.useBvh != 0) {
#if MRAD_SPLIT_SLAB > 1
		if (ray.fast && d.x != 0.0f && d.y != 0.0f && d.z != 0.0f)
			traceBvh<kMetrics, kCap, kShort, kSlabFastNz, kBatch, kTop>(S, R, ray, best, c, lds, top);
		else
#endif
		if (ray.fast) traceBvh<kMetrics, kCap, kShort, kSlabFast, kBatch, kTop>(S, R, ray, best, c, lds, top);
		else traceBvh<kMetrics, kCap, kShort, kSlabIeee, kBatch, kTop>(S, R, ray, best, c, lds, top);
	}
#else
	if (R.useBvh != 0) traceBvh<kMetrics, kCap, kShort, kSlabAny, kBatch, kTop>(S, R, ray, best, c, lds, top);
#endif
	else traceLinear<kMetrics>(S, R, ray, best, c);
	c.add(2, (uint32_t) R.planeCount);
	c.add(5, (uint32_t) R.planeCount);
	tracePlanes(S, R, o, d, best);
	return best.kind >= 0;
}

// reflect_dir / refract_dir / sampleUnitSphere (GpuKernelSources.java:1022-1059).
__device__ __forceinline__ f3 reflectDir(f3 n, f3 in) { return n * (2.0f * dot(in, n)) - in; }
__device__ inline f3 refractDir(f3 incoming, f3 n, float eta) {
	f3 i = normalizeSafe(incoming);
	float c1 = dot(i, n);
	float ri = (c1 >= 0.0f) ? eta : -1.0f / eta;
	float invRi2 = 1.0f / (ri * ri);
	float c2 = 1.0f - (1.0f - c1 * c1) * invRi2;
	if (c2 <= 0.0f) return normalizeSafe(reflectDir(n, i));
	float sq = sqrtf(c2);
	return normalizeSafe(n * (c1 - sq * ri) - i);
}
__device__ __forceinline__ f3 sampleUnitSphere(uint32_t &rng) {
	float z = rand01(rng) * 2.0f - 1.0f;
	float t = rand01(rng) * 2.0f * 3.141592654f;
	float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
	float sn, cs;
	sincos_cephes(t, sn, cs);
	return mk(r * cs, r * sn, z);
}

// One scattering event of tracePath (GpuKernelSources.java:1120-1161);
// returns false when the path is absorbed (all lobe weights zero).
__device__ __forceinline__ bool scatter(const MaterialF &m, f3 normal, f3 &dir, uint32_t &rng, f3 &weight) {
	f3 in = -dir;
	float wD = luminance(m.diffuse), wR = luminance(m.reflective), wT = luminance(m.refractive);
	float sum = wD + wR + wT;
	if (sum <= 0.0f) return false;
	float xi = rand01(rng) * sum;
	if (xi < wD) {
		float p = wD / sum;
		dir = sampleUnitSphere(rng) + normal;
		weight = divs(m.diffuse, p);
	} else if (xi < wD + wR) {
		float p = wR / sum;
		dir = reflectDir(normal, in);
		weight = divs(m.reflective, p);
	} else {
		float p = wT / sum;
		dir = refractDir(in, normal, m.ior);
		weight = divs(m.refractive, p);
	}
	return true;
}

// Russian roulette after bounce rrStart (no counterpart in the reference).
__device__ __forceinline__ bool roulette(int depth, int maxDepth, int rrStart, f3 &T, uint32_t &rng) {
	if (rrStart < 0 || depth + 1 < rrStart || depth + 1 >= maxDepth) return true;
	float m = fmaxf(T.x, fmaxf(T.y, T.z));
	float q = fminf(fmaxf(m, 0.05f), 0.95f);
	if (rand01(rng) >= q) return false;
	T = T * (1.0f / q);
	return true;
}

// One path segment after its closest hit has been resolved: emission,
// scattering, throughput and roulette.  Returns false when the path ends here.
__device__ __forceinline__ bool shadeResolved(const DevScene &S, const float *bgRGB, bool found, const HitDetail &h,
                                              int depth, int maxDepth, int rrStart, f3 &o, f3 &d, f3 &T, f3 &L,
                                              uint32_t &rng) {
	if (!found) {
		L = L + T * mk(bgRGB[0], bgRGB[1], bgRGB[2]);
		return false;
	}
	MaterialF m = loadMaterial(S, h.matIndex);
	L = L + T * m.emittance;
	f3 w;
	if (!scatter(m, h.normal, d, rng, w)) return false;
	o = h.pos + d * kEps;
	T = T * w;
	return roulette(depth, maxDepth, rrStart, T, rng);
}

// One path segment after its closest-hit query: emission, scattering,
// throughput and roulette.  Returns false when the path ends here.
__device__ __forceinline__ bool shadeSegment(const DevScene &S, const SceneRange &R, bool found, const Best &best,
                                             int depth, int maxDepth, int rrStart, f3 &o, f3 &d, f3 &T, f3 &L,
                                             uint32_t &rng) {
	if (!found) {
		L = L + T * mk(R.bg[0], R.bg[1], R.bg[2]);
		return false;
	}
	HitDetail h = resolveHit(S, o, d, best.t, best.kind, best.index);
	MaterialF m = loadMaterial(S, h.matIndex);
	L = L + T * m.emittance;
	f3 w;
	if (!scatter(m, h.normal, d, rng, w)) return false;
	o = h.pos + d * kEps;
	T = T * w;
	return roulette(depth, maxDepth, rrStart, T, rng);
}

// tracePath (GpuKernelSources.java:1061-1165).
template <bool kMetrics, int kCap, int kBatch = 0, int kTop = 0>
__device__ inline f3 tracePath(const DevScene &S, const SceneRange &R, f3 o, f3 d, int maxDepth, int rrStart,
                               uint32_t &rng, Counters<kMetrics> &c, LdsStack lds, TopNodes top = {}) {
	f3 T = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
	for (int depth = 0; depth < maxDepth; ++depth) {
		c.add(0, 1);
		Best best;
		const bool found = traceClosest<kMetrics, kCap, kShortStack, kBatch, kTop>(S, R, o, d, best, c, lds, top);
		if (!shadeSegment(S, R, found, best, depth, maxDepth, rrStart, o, d, T, L, rng)) break;
	}
	return L;
}

}  // namespace mrad
