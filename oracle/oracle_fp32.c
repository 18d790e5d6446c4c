/*
 * FP32 CPU restatement of the reference GPU path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is built into
 * oracle/liboracle.so and may be loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, always as
 * the checker and never as the thing measured or shipped.  The product path
 * (master-rad_amd/) never links or calls it.
 *
 * It follows src/xyz/marsavic/gfxlab/gpu/GpuKernelSources.java line by line,
 * over the reference's own packed wire format (GpuScene.java:24-31):
 *   spheres f32x5, boxes f32x7, planes f32x26, affine spheres/boxes f32x22,
 *   node bounds f32x6, node data i32x4, refs i32, roots i32, materials f32x13.
 * Every function cites the reference lines it restates.  It must be compiled
 * with -ffp-contract=off and without -ffast-math: only the fmaf calls the
 * reference writes are fused, "/" and sqrtf are IEEE correctly rounded, and
 * denormals are kept.
 *
 * Two deliberate, documented departures, applied identically in the HIP
 * product so both compute the same bits:
 *   - cosf/sinf of sampleUnitSphere (GpuKernelSources.java:1050-1055) use
 *     or_sincosf below: the single-precision Cody-Waite reduction and minimax
 *     polynomials of the Cephes math library (S. L. Moshier, sinf.c/cosf.c),
 *     evaluated with plain float operations.  The reference's libdevice bits
 *     are not reproducible anyway.
 *   - optional Russian roulette (absent from the reference): with
 *     rrStartDepth < 0 the path is exactly the reference's.
 *
 * Pinning: tests/test_oracle_*.py check every function here against closed
 * forms (axis-aligned and grazing rays, inside-origin exits, zero direction
 * components), brute-force FP64 intersection, BVH-versus-linear equality, the
 * exact counter identities of the reference (N_AABB = roots + 2 N_internal,
 * leaf-kind partition) and the Lambertian furnace closed form.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define EPSILON 1e-4f
#define REF_TYPE_MASK 0xC0000000u
#define REF_INDEX_MASK 0x3FFFFFFFu
#define REF_TYPE_SPHERE 0x00000000u
#define REF_TYPE_BOX 0x40000000u
#define REF_TYPE_AFFINE_SPHERE 0x80000000u
#define REF_TYPE_AFFINE_BOX 0xC0000000u
#define LEAF_MIXED -1
#define LEAF_SPHERE -2
#define LEAF_BOX -3
#define LEAF_AFFINE_SPHERE -4
#define LEAF_AFFINE_BOX -5
#define MAX_STACK 256

/* Counter slots, GpuKernelSources.java:30-43. */
enum {
	ST_RAYS = 0, ST_AABB, ST_PRIM, ST_SPHERE, ST_BOX, ST_PLANE, ST_ASPHERE,
	ST_ABOX, ST_OVERFLOW, ST_MAXSTACK, ST_INTERNAL, ST_LEAF, ST_HOMOG, ST_MIXED,
	ST_COUNT
};

typedef struct { float x, y, z; } v3;

typedef struct {
	const float *spheres; int sphereCount;
	const float *boxes; int boxCount;
	const float *planes; int planeCount;
	const float *affineSpheres; int affineSphereCount;
	const float *affineBoxes; int affineBoxCount;
	const float *nodeBounds; int nodeCount;
	const int *nodeData;
	const int *refs; int refCount;
	const int *roots; int rootCount;
	int useBvh;
	const float *materials; int materialCount;
	float bg[3];
	int stackSize;      /* BVH_STACK_SIZE, 1..256 */
} OrScene;

typedef struct { v3 diffuse, reflective, refractive, emittance; float ior; } Mat;

typedef struct {
	int hit; float t; int order; int matIndex;
	v3 pos, normal; Mat m;
} Hit;

/* ---- float3 helpers, GpuKernelSources.java:45-118 ---- */
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 vmuls(v3 v, float s) { return mk(v.x * s, v.y * s, v.z * s); }
static inline v3 vmul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdivs(v3 v, float s) { float inv = 1.0f / s; return mk(v.x * inv, v.y * inv, v.z * inv); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 normalizeSafe(v3 v) { float len = sqrtf(fmaxf(vdot(v, v), 1e-20f)); return vdivs(v, len); }

/* GpuKernelSources.java:101-110 */
static inline v3 normalizeTransformedNormal(v3 v) {
	float l2 = vdot(v, v);
	if (l2 >= 1.0e-20f && isfinite(l2)) return vdivs(v, sqrtf(l2));
	float scale = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
	if (!(scale > 0.0f) || !isfinite(scale)) return mk(0.0f, 0.0f, 0.0f);
	v3 s = mk(v.x / scale, v.y / scale, v.z / scale);
	return vdivs(s, sqrtf(vdot(s, s)));
}

/* ---- RNG, GpuKernelSources.java:165-179 ---- */
uint32_t or_hash_u32(uint32_t x) {
	x ^= x >> 17; x *= 0xed5ad4bbu;
	x ^= x >> 11; x *= 0xac4c1b51u;
	x ^= x >> 15; x *= 0x31848babu;
	x ^= x >> 14;
	return x;
}
static inline float rand01(uint32_t *s) {
	*s = or_hash_u32(*s + 0x9e3779b9u);
	return (float) (*s >> 8) * 0x1p-24f;
}
float or_rand01(uint32_t *s) { return rand01(s); }

/* luminance, GpuKernelSources.java:181-183 */
static inline float luminance(v3 c) { return c.x * 0.212655f + c.y * 0.715158f + c.z * 0.072187f; }

/* Cephes single-precision sine/cosine (Moshier), float arithmetic only. */
void or_sincosf(float xx, float *sOut, float *cOut) {
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
	float s = (j == 1 || j == 2) ? pc : ps;
	float c = (j == 1 || j == 2) ? ps : pc;
	*sOut = signS < 0 ? -s : s;
	*cOut = signC < 0 ? -c : c;
}

/* positiveModulo / material loads, GpuKernelSources.java:185-212 */
static inline float positiveModulo(float x, float y) { return x - y * floorf(x / y); }
static inline Mat loadMaterialByIndex(const float *materials, int index) {
	Mat m; memset(&m, 0, sizeof m);
	if (index < 0) return m;
	const float *s = materials + index * 13;
	m.diffuse = mk(s[0], s[1], s[2]);
	m.reflective = mk(s[3], s[4], s[5]);
	m.refractive = mk(s[6], s[7], s[8]);
	m.emittance = mk(s[9], s[10], s[11]);
	m.ior = s[12];
	return m;
}

/* Plane grid material, GpuKernelSources.java:214-269 */
static int choosePlaneMaterialIndex(const float *p, v3 hp) {
	const float *g = p + 18;
	int surface = (int) g[5];
	if (g[0] == 0.0f) return surface;
	v3 point = mk(p[4], p[5], p[6]);
	v3 e = mk(p[7], p[8], p[9]), f = mk(p[10], p[11], p[12]);
	float eLenSq = p[13], fLenSq = p[14], e_f = p[15], f_e = p[16], sinSq = p[17];
	v3 rel = vsub(hp, point);
	float b_e = vdot(rel, e) / eLenSq;
	float b_f = vdot(rel, f) / fLenSq;
	float u = (b_e - b_f * f_e) / sinSq;
	float v = (b_f - b_e * e_f) / sinSq;
	float hx = g[3] * 0.5f, hy = g[4] * 0.5f;
	float modX = positiveModulo(u + hx, g[1]);
	float modY = positiveModulo(v + hy, g[2]);
	if (modX < g[3] || modY < g[4]) return (int) g[7];
	return (int) g[6];
}

/* GpuKernelSources.java:276-309 */
static inline float affineRayShift(const float *d, v3 o, v3 dir) {
	v3 c = mk(d[3], d[7], d[11]);
	float den = vdot(dir, dir);
	return den > 0.0f ? fmaxf(0.0f, vdot(vsub(c, o), dir) / den) : 0.0f;
}
static inline v3 applyAffineShiftedPoint(const float *d, v3 o, v3 dir, float shift) {
	v3 r = mk(fmaf(dir.x, shift, o.x - d[3]), fmaf(dir.y, shift, o.y - d[7]), fmaf(dir.z, shift, o.z - d[11]));
	return mk(d[0] * r.x + d[1] * r.y + d[2] * r.z,
	          d[4] * r.x + d[5] * r.y + d[6] * r.z,
	          d[8] * r.x + d[9] * r.y + d[10] * r.z);
}
static inline v3 applyAffineVector(const float *d, v3 v) {
	return mk(d[0] * v.x + d[1] * v.y + d[2] * v.z,
	          d[4] * v.x + d[5] * v.y + d[6] * v.z,
	          d[8] * v.x + d[9] * v.y + d[10] * v.z);
}
static inline v3 applyLinear9(const float *d, v3 v) {
	return mk(d[0] * v.x + d[1] * v.y + d[2] * v.z,
	          d[3] * v.x + d[4] * v.y + d[5] * v.z,
	          d[6] * v.x + d[7] * v.y + d[8] * v.z);
}

/* GpuKernelSources.java:317-336 */
static int normalizedDirection(v3 d, v3 *unit, float *invLen) {
	float l2 = vdot(d, d);
	if (isfinite(l2) && fabsf(l2 - 1.0f) <= 2.0e-6f) { *unit = d; *invLen = 1.0f; return 1; }
	float scale = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
	if (!(scale > 0.0f) || !isfinite(scale)) return 0;
	v3 s = mk(d.x / scale, d.y / scale, d.z / scale);
	float sl = sqrtf(vdot(s, s));
	float dl = scale * sl;
	float inv = 1.0f / dl;
	if (!(inv > 0.0f) || !isfinite(inv)) return 0;
	*unit = vmuls(s, 1.0f / sl);
	*invLen = inv;
	return 1;
}

/* traceSphere, GpuKernelSources.java:338-399 */
static int traceSphere(const float *sp, int count, const float *mats, v3 o, v3 d, int orderBase, Hit *hit) {
	int found = 0;
	float closestT = hit->t;
	for (int i = 0; i < count; ++i) {
		const float *s = sp + i * 5;
		v3 c = mk(s[0], s[1], s[2]);
		float radius = s[3];
		int mi = (int) s[4];
		v3 oc = vsub(o, c);
		float a = vdot(d, d);
		float b = vdot(oc, d);
		float centerT = -b / a;
		v3 cl = mk(fmaf(d.x, centerT, oc.x), fmaf(d.y, centerT, oc.y), fmaf(d.z, centerT, oc.z));
		float corr = -vdot(cl, d) / a;
		centerT += corr;
		cl = mk(fmaf(d.x, corr, cl.x), fmaf(d.y, corr, cl.y), fmaf(d.z, corr, cl.z));
		float radialSquared = fmaf(radius, radius, -vdot(cl, cl));
		if (radialSquared <= 0.0f) continue;
		float halfSpan = sqrtf(radialSquared / a);
		float t1 = centerT - halfSpan, t2 = centerT + halfSpan;
		float hitT, radialOffset;
		if (t1 > EPSILON) { hitT = t1; radialOffset = -halfSpan; }
		else if (t2 > EPSILON) { hitT = t2; radialOffset = halfSpan; }
		else continue;
		if (hitT < closestT) {
			closestT = hitT;
			hit->hit = 1;
			hit->t = hitT;
			hit->pos = vadd(o, vmuls(d, hitT));
			v3 radial = mk(fmaf(d.x, radialOffset, cl.x), fmaf(d.y, radialOffset, cl.y), fmaf(d.z, radialOffset, cl.z));
			hit->normal = normalizeSafe(radial);
			hit->m = loadMaterialByIndex(mats, mi);
			hit->matIndex = mi;
			hit->order = orderBase + i;
			found = 1;
		}
	}
	return found;
}

/* traceAffineSphere, GpuKernelSources.java:401-461 */
static int traceAffineSphere(const float *as, int count, const float *mats, v3 o, v3 d, int orderBase, Hit *hit) {
	int found = 0;
	float closestT = hit->t;
	for (int i = 0; i < count; ++i) {
		const float *a = as + i * 22;
		float shift = affineRayShift(a, o, d);
		v3 lo = applyAffineShiftedPoint(a, o, d, shift);
		v3 ld = applyAffineVector(a, d);
		float afterLocal = EPSILON - shift;
		float solveA = vdot(ld, ld);
		v3 solveDir = ld;
		float pscale = 1.0f;
		if (!(solveA > 0.0f) || !isfinite(solveA)) {
			v3 u; float inv;
			if (!normalizedDirection(ld, &u, &inv)) continue;
			solveDir = u; solveA = 1.0f; pscale = inv;
		}
		float scaledAfterT = afterLocal / pscale;
		float centerT = -vdot(lo, solveDir) / solveA;
		v3 cl = vadd(lo, vmuls(solveDir, centerT));
		float rs = fmaf(-cl.x, cl.x, fmaf(-cl.y, cl.y, fmaf(-cl.z, cl.z, 1.0f)));
		if (rs <= 0.0f) continue;
		float halfSpan = sqrtf(rs / solveA);
		float t1 = centerT - halfSpan, t2 = centerT + halfSpan;
		float sh;
		if (t1 > scaledAfterT) sh = t1;
		else if (t2 > scaledAfterT) sh = t2;
		else continue;
		float hitT = sh * pscale;
		float worldT = shift + hitT;
		if (worldT >= closestT) continue;
		v3 lh = vadd(lo, vmuls(ld, hitT));
		v3 wn = normalizeTransformedNormal(applyLinear9(a + 12, lh));
		int mi = (int) a[21];
		hit->hit = 1; hit->t = worldT;
		hit->pos = vadd(o, vmuls(d, worldT));
		hit->normal = wn;
		hit->m = loadMaterialByIndex(mats, mi);
		hit->matIndex = mi;
		hit->order = orderBase + i;
		closestT = worldT;
		found = 1;
	}
	return found;
}

/* Slab of one axis for the unit box / primitive box.  Returns 0 when the
 * zero-direction branch rejects the primitive (GpuKernelSources.java:476-515,600-639). */
static inline int slabAxis(float oc, float dc, float lo, float hi, float *tmin, float *tmax) {
	if (dc == 0.0f) {
		if (oc < lo || oc > hi) return 0;
		*tmin = -INFINITY; *tmax = INFINITY;
		return 1;
	}
	float t1 = (lo - oc) / dc, t2 = (hi - oc) / dc;
	*tmin = fminf(t1, t2); *tmax = fmaxf(t1, t2);
	return 1;
}

/* Shared enter/exit axis selection, GpuKernelSources.java:517-556 and :641-680. */
static inline int boxCandidate(float txMin, float txMax, float tyMin, float tyMax, float tzMin, float tzMax,
		float after, v3 d, float *candT, int *axis, float *sign) {
	float tEnter = fmaxf(fmaxf(txMin, tyMin), tzMin);
	float tExit = fminf(fminf(txMax, tyMax), tzMax);
	if (tEnter >= tExit || tExit <= after) return 0;
	int axisEnter = 0; float ev = txMin;
	if (tyMin > ev) { ev = tyMin; axisEnter = 1; }
	if (tzMin > ev) axisEnter = 2;
	int axisExit = 0; float xv = txMax;
	if (tyMax < xv) { xv = tyMax; axisExit = 1; }
	if (tzMax < xv) axisExit = 2;
	if (tEnter > after) {
		*candT = tEnter; *axis = axisEnter;
		float dc = axisEnter == 0 ? d.x : (axisEnter == 1 ? d.y : d.z);
		*sign = dc > 0.0f ? -1.0f : 1.0f;
	} else {
		*candT = tExit; *axis = axisExit;
		float dc = axisExit == 0 ? d.x : (axisExit == 1 ? d.y : d.z);
		*sign = dc >= 0.0f ? 1.0f : -1.0f;
	}
	return 1;
}
static inline v3 axisNormal(int axis, float sign) {
	return axis == 0 ? mk(sign, 0.0f, 0.0f) : (axis == 1 ? mk(0.0f, sign, 0.0f) : mk(0.0f, 0.0f, sign));
}

/* traceAffineBox, GpuKernelSources.java:463-586 */
static int traceAffineBox(const float *ab, int count, const float *mats, v3 o, v3 d, int orderBase, Hit *hit) {
	int found = 0;
	float closestT = hit->t;
	for (int i = 0; i < count; ++i) {
		const float *a = ab + i * 22;
		float shift = affineRayShift(a, o, d);
		v3 lo = applyAffineShiftedPoint(a, o, d, shift);
		v3 ld = applyAffineVector(a, d);
		float afterLocal = EPSILON - shift;
		float x0, x1, y0, y1, z0, z1;
		if (!slabAxis(lo.x, ld.x, -1.0f, 1.0f, &x0, &x1)) continue;
		if (!slabAxis(lo.y, ld.y, -1.0f, 1.0f, &y0, &y1)) continue;
		if (!slabAxis(lo.z, ld.z, -1.0f, 1.0f, &z0, &z1)) continue;
		float ct; int axis; float sign;
		if (!boxCandidate(x0, x1, y0, y1, z0, z1, afterLocal, ld, &ct, &axis, &sign)) continue;
		float worldT = shift + ct;
		if (worldT >= closestT) continue;
		v3 wn = normalizeTransformedNormal(applyLinear9(a + 12, axisNormal(axis, sign)));
		int mi = (int) a[21];
		hit->hit = 1; hit->t = worldT;
		hit->pos = vadd(o, vmuls(d, worldT));
		hit->normal = wn;
		hit->m = loadMaterialByIndex(mats, mi);
		hit->matIndex = mi;
		hit->order = orderBase + i;
		closestT = worldT;
		found = 1;
	}
	return found;
}

/* traceBox, GpuKernelSources.java:588-706 */
static int traceBox(const float *bx, int count, const float *mats, v3 o, v3 d, int orderBase, Hit *hit) {
	int found = 0;
	float closestT = hit->t;
	for (int i = 0; i < count; ++i) {
		const float *b = bx + i * 7;
		int mi = (int) b[6];
		float x0, x1, y0, y1, z0, z1;
		if (!slabAxis(o.x, d.x, b[0], b[3], &x0, &x1)) continue;
		if (!slabAxis(o.y, d.y, b[1], b[4], &y0, &y1)) continue;
		if (!slabAxis(o.z, d.z, b[2], b[5], &z0, &z1)) continue;
		float ct; int axis; float sign;
		if (!boxCandidate(x0, x1, y0, y1, z0, z1, EPSILON, d, &ct, &axis, &sign)) continue;
		if (ct >= closestT) continue;
		hit->hit = 1; hit->t = ct;
		hit->pos = vadd(o, vmuls(d, ct));
		hit->normal = axisNormal(axis, sign);
		hit->m = loadMaterialByIndex(mats, mi);
		hit->matIndex = mi;
		hit->order = orderBase + i;
		closestT = ct;
		found = 1;
	}
	return found;
}

/* tracePlane, GpuKernelSources.java:708-734 */
static int tracePlane(const float *pl, int count, const float *mats, v3 o, v3 d, int orderBase, Hit *hit) {
	int found = 0;
	float closestT = hit->t;
	for (int i = 0; i < count; ++i) {
		const float *p = pl + i * 26;
		v3 n = mk(p[0], p[1], p[2]);
		float denom = vdot(n, d);
		if (denom == 0.0f) continue;
		float t = -(vdot(n, o) + p[3]) / denom;
		if (t <= EPSILON || t >= closestT) continue;
		v3 hp = vadd(o, vmuls(d, t));
		hit->hit = 1; hit->t = t; hit->pos = hp; hit->normal = n;
		hit->matIndex = choosePlaneMaterialIndex(p, hp);
		hit->m = loadMaterialByIndex(mats, hit->matIndex);
		hit->order = orderBase + i;
		closestT = t;
		found = 1;
	}
	return found;
}

/* bvhEntryDistance, GpuKernelSources.java:736-770 */
float or_bvh_entry_distance(const float *bounds, int node, v3 o, v3 d, float beforeT) {
	const float *b = bounds + node * 6;
	float tMin = EPSILON;
	float cb = isfinite(beforeT) ? fmaf(beforeT, 1.0000015f, 1.0e-6f) : beforeT;
	float tMax = cb;
	for (int axis = 0; axis < 3; axis++) {
		float oc = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
		float dc = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
		float lo = b[axis], hi = b[axis + 3];
		if (dc == 0.0f) {
			if (oc < lo || oc > hi) return INFINITY;
			continue;
		}
		float t0 = (lo - oc) / dc, t1 = (hi - oc) / dc;
		if (t0 > t1) { float tmp = t0; t0 = t1; t1 = tmp; }
		if (t1 > 0.0f) t1 *= 1.00000072f;
		tMin = fmaxf(tMin, t0);
		tMax = fminf(tMax, t1);
		if (tMax < tMin) return INFINITY;
	}
	return tMin <= cb ? tMin : INFINITY;
}

static inline int typeSlot(uint32_t type) {
	return type == REF_TYPE_BOX ? 1 : type == REF_TYPE_AFFINE_SPHERE ? 3 : type == REF_TYPE_AFFINE_BOX ? 4 : 0;
}

/* tracePrimitiveRef, GpuKernelSources.java:772-806 */
static int tracePrimitiveRef(const OrScene *s, uint32_t ref, uint32_t type, v3 o, v3 d, Hit *hit) {
	int index = (int) (ref & REF_INDEX_MASK);
	Hit cand; memset(&cand, 0, sizeof cand);
	cand.t = nextafterf(hit->t, INFINITY);
	int found, order;
	if (type == REF_TYPE_BOX) {
		order = s->sphereCount + index;
		found = traceBox(s->boxes + index * 7, 1, s->materials, o, d, order, &cand);
	} else if (type == REF_TYPE_AFFINE_SPHERE) {
		order = s->sphereCount + s->boxCount + index;
		found = traceAffineSphere(s->affineSpheres + index * 22, 1, s->materials, o, d, order, &cand);
	} else if (type == REF_TYPE_AFFINE_BOX) {
		order = s->sphereCount + s->boxCount + s->affineSphereCount + index;
		found = traceAffineBox(s->affineBoxes + index * 22, 1, s->materials, o, d, order, &cand);
	} else {
		order = index;
		found = traceSphere(s->spheres + index * 5, 1, s->materials, o, d, order, &cand);
	}
	if (!found || cand.t > hit->t || (cand.t == hit->t && order >= hit->order)) return 0;
	*hit = cand;
	hit->order = order;
	return 1;
}

/* traceBvhLeaf, GpuKernelSources.java:808-841 */
static int traceBvhLeaf(const OrScene *s, int kind, int first, int count, v3 o, v3 d, Hit *hit, uint64_t *st) {
	if (st) st[ST_PRIM] += (uint64_t) count;
	int found = 0;
	int leafType = kind == LEAF_SPHERE ? (int) REF_TYPE_SPHERE
		: kind == LEAF_BOX ? (int) REF_TYPE_BOX
		: kind == LEAF_AFFINE_SPHERE ? (int) REF_TYPE_AFFINE_SPHERE
		: kind == LEAF_AFFINE_BOX ? (int) REF_TYPE_AFFINE_BOX : LEAF_MIXED;
	for (int i = 0; i < count; i++) {
		int ri = first + i;
		if (ri < 0 || ri >= s->refCount) continue;
		uint32_t ref = (uint32_t) s->refs[ri];
		uint32_t type = leafType == LEAF_MIXED ? (ref & REF_TYPE_MASK) : (uint32_t) leafType;
		if (st) st[ST_SPHERE + typeSlot(type)]++;
		found |= tracePrimitiveRef(s, ref, type, o, d, hit);
	}
	return found;
}

static inline void bumpMax(uint64_t *st, int size) {
	if (st && (uint64_t) size > st[ST_MAXSTACK]) st[ST_MAXSTACK] = (uint64_t) size;
}

/* traceBvh, GpuKernelSources.java:844-961 */
static int traceBvh(const OrScene *s, v3 o, v3 d, Hit *hit, uint64_t *st) {
	if (s->nodeCount <= 0 || s->refCount <= 0 || s->rootCount <= 0 || s->rootCount > 4) return 0;
	const int cap = s->stackSize;
	int nodeStack[MAX_STACK];
	float entryStack[MAX_STACK];
	int size = 0, found = 0, overflow = 0;
	for (int r = 0; r < s->rootCount; r++) {
		int ri = s->roots[r];
		if (ri < 0 || ri >= s->nodeCount) { overflow = 1; break; }
		if (st) st[ST_AABB]++;
		float e = or_bvh_entry_distance(s->nodeBounds, ri, o, d, hit->t);
		if (!isfinite(e)) continue;
		if (size >= cap) { overflow = 1; break; }
		int pos = size;
		while (pos > 0 && e >= entryStack[pos - 1]) {
			nodeStack[pos] = nodeStack[pos - 1];
			entryStack[pos] = entryStack[pos - 1];
			pos--;
		}
		nodeStack[pos] = ri; entryStack[pos] = e;
		size++;
		bumpMax(st, size);
	}
	while (!overflow && size > 0) {
		int node = nodeStack[--size];
		float entry = entryStack[size];
		float ch = isfinite(hit->t) ? fmaf(hit->t, 1.0000015f, 1.0e-6f) : hit->t;
		if (entry > ch) continue;
		const int *nd = s->nodeData + node * 4;
		int leftOrKind = nd[0], right = nd[1], first = nd[2], count = nd[3];
		if (leftOrKind < 0) {
			if (st) {
				st[ST_LEAF]++;
				if (leftOrKind == LEAF_MIXED) st[ST_MIXED]++; else st[ST_HOMOG]++;
			}
			found |= traceBvhLeaf(s, leftOrKind, first, count, o, d, hit, st);
			continue;
		}
		if (st) st[ST_INTERNAL]++;
		float le = or_bvh_entry_distance(s->nodeBounds, leftOrKind, o, d, hit->t);
		float re = or_bvh_entry_distance(s->nodeBounds, right, o, d, hit->t);
		if (st) st[ST_AABB] += 2;
		int nearN = leftOrKind, farN = right;
		float nearE = le, farE = re;
		if (re < le) { nearN = right; farN = leftOrKind; nearE = re; farE = le; }
		if (isfinite(farE)) {
			if (size >= cap) { overflow = 1; break; }
			nodeStack[size] = farN; entryStack[size++] = farE;
			bumpMax(st, size);
		}
		if (isfinite(nearE)) {
			if (size >= cap) { overflow = 1; break; }
			nodeStack[size] = nearN; entryStack[size++] = nearE;
			bumpMax(st, size);
		}
	}
	if (overflow) {
		if (st) st[ST_OVERFLOW]++;
		for (int i = 0; i < s->refCount; i++) {
			uint32_t ref = (uint32_t) s->refs[i];
			if (st) { st[ST_PRIM]++; st[ST_SPHERE + typeSlot(ref & REF_TYPE_MASK)]++; }
			found |= tracePrimitiveRef(s, ref, ref & REF_TYPE_MASK, o, d, hit);
		}
	}
	return found;
}

/* traceClosest, GpuKernelSources.java:963-1020 */
static int traceClosest(const OrScene *s, v3 o, v3 d, Hit *hit, uint64_t *st) {
	memset(hit, 0, sizeof *hit);
	hit->hit = 0; hit->t = 1e30f; hit->order = 0x7FFFFFFF;
	int found = 0;
	if (s->useBvh != 0 && s->nodeCount > 0) {
		found = traceBvh(s, o, d, hit, st);
	} else {
		if (st) {
			st[ST_PRIM] += (uint64_t) s->sphereCount + s->boxCount + s->affineSphereCount + s->affineBoxCount;
			st[ST_SPHERE] += s->sphereCount; st[ST_BOX] += s->boxCount;
			st[ST_ASPHERE] += s->affineSphereCount; st[ST_ABOX] += s->affineBoxCount;
		}
		int bb = s->sphereCount, asb = bb + s->boxCount, abb = asb + s->affineSphereCount;
		found = traceSphere(s->spheres, s->sphereCount, s->materials, o, d, 0, hit);
		found |= traceBox(s->boxes, s->boxCount, s->materials, o, d, bb, hit);
		found |= traceAffineSphere(s->affineSpheres, s->affineSphereCount, s->materials, o, d, asb, hit);
		found |= traceAffineBox(s->affineBoxes, s->affineBoxCount, s->materials, o, d, abb, hit);
	}
	if (st) { st[ST_PRIM] += (uint64_t) s->planeCount; st[ST_PLANE] += s->planeCount; }
	int pb = s->sphereCount + s->boxCount + s->affineSphereCount + s->affineBoxCount;
	found |= tracePlane(s->planes, s->planeCount, s->materials, o, d, pb, hit);
	return found;
}

/* GpuKernelSources.java:1022-1059 */
static inline v3 reflect_dir(v3 n, v3 in) { return vsub(vmuls(n, 2.0f * vdot(in, n)), in); }
static int refract_dir(v3 incoming, v3 n, float eta, v3 *res) {
	v3 i = normalizeSafe(incoming);
	float c1 = vdot(i, n);
	float ri = (c1 >= 0.0f) ? eta : -1.0f / eta;
	float invRi2 = 1.0f / (ri * ri);
	float c2 = 1.0f - (1.0f - c1 * c1) * invRi2;
	if (c2 <= 0.0f) { *res = normalizeSafe(reflect_dir(n, i)); return 0; }
	float sq = sqrtf(c2);
	*res = normalizeSafe(vsub(vmuls(n, c1 - sq * ri), i));
	return 1;
}
static inline v3 sampleUnitSphere(uint32_t *rng) {
	float z = rand01(rng) * 2.0f - 1.0f;
	float t = rand01(rng) * 2.0f * 3.141592654f;
	float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
	float sn, cs;
	or_sincosf(t, &sn, &cs);
	return mk(r * cs, r * sn, z);
}

/* tracePath, GpuKernelSources.java:1061-1165, plus optional Russian roulette
 * after bounce rrStart (survival q = clamp(max(T), 0.05, 0.95)). */
static v3 tracePath(const OrScene *s, v3 o, v3 d, int maxDepth, int rrStart, uint32_t *rng, uint64_t *st) {
	v3 T = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
	v3 bg = mk(s->bg[0], s->bg[1], s->bg[2]);
	for (int depth = 0; depth < maxDepth; ++depth) {
		if (st) st[ST_RAYS]++;
		Hit h;
		int hf = traceClosest(s, o, d, &h, st);
		if (!hf) { L = vadd(L, vmul(T, bg)); break; }
		L = vadd(L, vmul(T, h.m.emittance));
		v3 n = h.normal, in = vneg(d);
		v3 cD = h.m.diffuse, cR = h.m.reflective, cT = h.m.refractive;
		float wD = luminance(cD), wR = luminance(cR), wT = luminance(cT);
		float sum = wD + wR + wT;
		if (sum <= 0.0f) break;
		float xi = rand01(rng) * sum;
		v3 nd, wc;
		float p;
		if (xi < wD) {
			p = wD / sum;
			nd = vadd(sampleUnitSphere(rng), n);
			wc = vdivs(cD, p);
		} else if (xi < wD + wR) {
			p = wR / sum;
			nd = reflect_dir(n, in);
			wc = vdivs(cR, p);
		} else {
			p = wT / sum;
			v3 rd;
			refract_dir(in, n, h.m.ior, &rd);
			nd = rd;
			wc = vdivs(cT, p);
		}
		o = vadd(h.pos, vmuls(nd, EPSILON));
		T = vmul(T, wc);
		d = nd;
		if (rrStart >= 0 && depth + 1 >= rrStart && depth + 1 < maxDepth) {
			float m = fmaxf(T.x, fmaxf(T.y, T.z));
			float q = fminf(fmaxf(m, 0.05f), 0.95f);
			if (rand01(rng) >= q) break;
			T = vmuls(T, 1.0f / q);
		}
	}
	return L;
}

/* Seed derivation of renderKernel, GpuKernelSources.java:1231-1253. */
uint32_t or_path_seed(uint64_t frameSeed, int frameIndex, uint32_t pixelIndex, int sampleOffset) {
	uint32_t base = or_hash_u32((uint32_t) (frameSeed & 0xFFFFFFFFull));
	base ^= or_hash_u32((uint32_t) (frameSeed >> 32));
	base ^= or_hash_u32((uint32_t) frameIndex);
	base ^= or_hash_u32(pixelIndex);
	base = or_hash_u32(base);
	uint32_t seed = base ^ or_hash_u32((uint32_t) sampleOffset);
	seed ^= or_hash_u32((uint32_t) sampleOffset * 0x9E3779B9u);
	return or_hash_u32(seed);
}

/* One path sample of pixel (x,y): renderKernel body, GpuKernelSources.java:1217-1282. */
/* Camera ray of renderKernel, GpuKernelSources.java:1256-1264: two jitter
 * draws from the path's RNG state, the pixel-centre offset of 0.5, the
 * aspect-scaled u and the downward v, and the unnormalized direction
 * fwd + right*u + up*v.  ray6 = origin.xyz, direction.xyz. */
void or_camera_ray(const float *cam, int width, int height, int x, int y, uint32_t *rng, float *ray6) {
	float jx = rand01(rng) - 0.5f;
	float jy = rand01(rng) - 0.5f;
	float aspect = (float) width / (float) height;
	float u = ((((float) x + 0.5f) + jx) * (2.0f / (float) width) - 1.0f) * aspect;
	float v = 1.0f - ((((float) y + 0.5f) + jy) * (2.0f / (float) height));
	v3 fwd = mk(cam[3], cam[4], cam[5]);
	v3 right = mk(cam[6], cam[7], cam[8]), up = mk(cam[9], cam[10], cam[11]);
	v3 dir = vadd(vadd(fwd, vmuls(right, u)), vmuls(up, v));
	ray6[0] = cam[0]; ray6[1] = cam[1]; ray6[2] = cam[2];
	ray6[3] = dir.x; ray6[4] = dir.y; ray6[5] = dir.z;
}

/* The camera ray of path (pixel, sample): seed, then or_camera_ray. */
static void primaryRay(const float *cam, int width, int height, int x, int y, int frameIndex, uint64_t frameSeed,
		int sampleOffset, uint32_t *rng, v3 *o, v3 *d) {
	uint32_t pixel = (uint32_t) y * (uint32_t) width + (uint32_t) x;
	float r[6];
	*rng = or_path_seed(frameSeed, frameIndex, pixel, sampleOffset);
	or_camera_ray(cam, width, height, x, y, rng, r);
	*o = mk(r[0], r[1], r[2]);
	*d = mk(r[3], r[4], r[5]);
}

static v3 renderPathSample(const OrScene *s, const float *cam, int width, int height, int x, int y,
		int maxDepth, int rrStart, int frameIndex, uint64_t frameSeed, int sampleOffset, uint64_t *st) {
	uint32_t rng;
	v3 pos, dir;
	primaryRay(cam, width, height, x, y, frameIndex, frameSeed, sampleOffset, &rng, &pos, &dir);
	return tracePath(s, pos, dir, maxDepth, rrStart, &rng, st);
}

/* Renders the listed pixels for samples 0..spp-1 with the kernel's float
 * accumulation and last-sample normalization (GpuKernelSources.java:1284-1293).
 * pixels == NULL means every pixel in row-major order. */
typedef struct {
	const OrScene *s; const float *cam; int width, height, maxDepth, rrStart, frameIndex, spp;
	uint64_t frameSeed; const int *pixels; int n; float *out; uint64_t *stats;
	int next; pthread_mutex_t lock;
} RenderJob;

static void renderRange(RenderJob *job, int begin, int end, uint64_t *st) {
	for (int k = begin; k < end; k++) {
		int p = job->pixels ? job->pixels[k] : k;
		int x = p % job->width, y = p / job->width;
		v3 acc = mk(0.0f, 0.0f, 0.0f);
		for (int so = 0; so < job->spp; so++) {
			v3 c = renderPathSample(job->s, job->cam, job->width, job->height, x, y, job->maxDepth,
					job->rrStart, job->frameIndex, job->frameSeed, so, st);
			acc = so == 0 ? c : vadd(acc, c);
			if (so == job->spp - 1) acc = vmuls(acc, 1.0f / (float) job->spp);
		}
		job->out[3 * k] = acc.x; job->out[3 * k + 1] = acc.y; job->out[3 * k + 2] = acc.z;
	}
}

static void *renderWorker(void *arg) {
	RenderJob *job = (RenderJob *) arg;
	uint64_t st[ST_COUNT] = {0};
	for (;;) {
		pthread_mutex_lock(&job->lock);
		int b = job->next; job->next += 64;
		pthread_mutex_unlock(&job->lock);
		if (b >= job->n) break;
		int e = b + 64 < job->n ? b + 64 : job->n;
		renderRange(job, b, e, job->stats ? st : NULL);
	}
	if (job->stats) {
		pthread_mutex_lock(&job->lock);
		for (int i = 0; i < ST_COUNT; i++) {
			if (i == ST_MAXSTACK) { if (st[i] > job->stats[i]) job->stats[i] = st[i]; }
			else job->stats[i] += st[i];
		}
		pthread_mutex_unlock(&job->lock);
	}
	return NULL;
}

int or_render_pixels(const OrScene *s, const float *cam, int width, int height, int maxDepth, int rrStart,
		int frameIndex, uint64_t frameSeed, int spp, const int *pixels, int n, float *out,
		uint64_t *stats, int threads) {
	if (!s || !cam || !out || width < 1 || height < 1 || spp < 1 || n < 0) return -1;
	if (s->stackSize < 1 || s->stackSize > MAX_STACK) return -1;
	if (stats) memset(stats, 0, sizeof(uint64_t) * ST_COUNT);
	RenderJob job = {s, cam, width, height, maxDepth, rrStart, frameIndex, spp, frameSeed, pixels, n, out, stats, 0};
	pthread_mutex_init(&job.lock, NULL);
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	pthread_t tid[256];
	for (int i = 1; i < threads; i++) pthread_create(&tid[i], NULL, renderWorker, &job);
	renderWorker(&job);
	for (int i = 1; i < threads; i++) pthread_join(tid[i], NULL);
	pthread_mutex_destroy(&job.lock);
	return 0;
}

/* Reference global order of the closest hit (GpuKernelSources.java:784-799, :1017). */
static int hitOrder(const Hit *h) { return h->order; }

/* Path-replay adjoint of one path (no counterpart in the reference).  The
 * path is re-traced with tracePath's arithmetic and random numbers; its
 * vertices are kept (throughput T_i, lobe weight w_i, diffuse-albedo factor
 * g_i = 1/(p_lobe q_roulette) or 0 for a non-diffuse lobe, emission Le_i,
 * primitive order) and swept backwards:
 *   dLoss/dLe_i     += delta * T_i
 *   dLoss/dalbedo_i += delta * T_i * g_i * Lin_i
 *   Lin_{i-1}        = Le_i + w_i * Lin_i
 * with Lin of the last vertex = background if its scattered ray missed, else 0.
 * Lobe-selection and roulette probabilities are treated as constants.  Sums in
 * double.  Returns the path radiance (float, as tracePath). */
static v3 adjointPath(const OrScene *s, v3 o, v3 d, int maxDepth, int rrStart, uint32_t *rng, const double *delta,
		double *gAlb, double *gEmit) {
	enum { MAXV = 32 };
	v3 Tv[MAXV], Wv[MAXV], Lev[MAXV];
	double gv[MAXV];
	int prim[MAXV];
	int n = 0;
	v3 T = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
	v3 bg = mk(s->bg[0], s->bg[1], s->bg[2]);
	v3 tail = mk(0.0f, 0.0f, 0.0f);
	for (int depth = 0; depth < maxDepth; ++depth) {
		Hit h;
		int hf = traceClosest(s, o, d, &h, NULL);
		if (!hf) { L = vadd(L, vmul(T, bg)); tail = bg; break; }
		L = vadd(L, vmul(T, h.m.emittance));
		Tv[n] = T; Lev[n] = h.m.emittance; prim[n] = hitOrder(&h);
		Wv[n] = mk(0.0f, 0.0f, 0.0f); gv[n] = 0.0;
		n++;
		v3 nrm = h.normal, in = vneg(d);
		v3 cD = h.m.diffuse, cR = h.m.reflective, cT = h.m.refractive;
		float wD = luminance(cD), wR = luminance(cR), wT = luminance(cT);
		float sum = wD + wR + wT;
		if (sum <= 0.0f) break;
		float xi = rand01(rng) * sum;
		v3 nd, wc;
		float p;
		int diffuse = 0;
		if (xi < wD) {
			p = wD / sum; diffuse = 1;
			nd = vadd(sampleUnitSphere(rng), nrm);
			wc = vdivs(cD, p);
		} else if (xi < wD + wR) {
			p = wR / sum;
			nd = reflect_dir(nrm, in);
			wc = vdivs(cR, p);
		} else {
			p = wT / sum;
			v3 rd;
			refract_dir(in, nrm, h.m.ior, &rd);
			nd = rd;
			wc = vdivs(cT, p);
		}
		o = vadd(h.pos, vmuls(nd, EPSILON));
		T = vmul(T, wc);
		d = nd;
		double scale = diffuse ? 1.0 / (double) p : 0.0;
		double q = 1.0;
		if (rrStart >= 0 && depth + 1 >= rrStart && depth + 1 < maxDepth) {
			float m = fmaxf(T.x, fmaxf(T.y, T.z));
			float qf = fminf(fmaxf(m, 0.05f), 0.95f);
			if (rand01(rng) >= qf) { Wv[n - 1] = wc; gv[n - 1] = scale; break; }
			T = vmuls(T, 1.0f / qf);
			q = qf;
		}
		Wv[n - 1] = mk((float) (wc.x / q), (float) (wc.y / q), (float) (wc.z / q));
		gv[n - 1] = scale / q;
		if (depth + 1 == maxDepth) break;
	}
	double lin[3] = {tail.x, tail.y, tail.z};
	/* a path killed by roulette or ended by depth after its last scatter has no incoming radiance */
	for (int i = n - 1; i >= 0; i--) {
		const double t[3] = {Tv[i].x, Tv[i].y, Tv[i].z};
		const double le[3] = {Lev[i].x, Lev[i].y, Lev[i].z};
		const double w[3] = {Wv[i].x, Wv[i].y, Wv[i].z};
		for (int c = 0; c < 3; c++) {
			gEmit[3 * prim[i] + c] += delta[c] * t[c];
			gAlb[3 * prim[i] + c] += delta[c] * t[c] * gv[i] * lin[c];
			lin[c] = le[c] + w[c] * lin[c];
		}
	}
	return L;
}

/* Adjoint of the whole frame: delta = dLoss/dImage (W*H*3 floats); gradients
 * per primitive in the reference's global order (grad arrays nPrim x 3).
 * Rows are split over `threads` threads, each summing into its own arrays;
 * the per-thread sums are added in thread order. */
typedef struct {
	const OrScene *s; const float *cam; int width, height, maxDepth, rrStart, frameIndex, spp; uint64_t frameSeed;
	const float *gradImage; double *gAlb, *gEmit; int nextRow; pthread_mutex_t lock;
} AdjointJob;

static void adjointRows(AdjointJob *j, double *gAlb, double *gEmit) {
	for (;;) {
		pthread_mutex_lock(&j->lock);
		int y = j->nextRow++;
		pthread_mutex_unlock(&j->lock);
		if (y >= j->height) break;
		for (int x = 0; x < j->width; x++) {
			uint32_t pixel = (uint32_t) y * (uint32_t) j->width + (uint32_t) x;
			double delta[3];
			for (int c = 0; c < 3; c++) delta[c] = (double) j->gradImage[3 * pixel + c] / (double) j->spp;
			for (int so = 0; so < j->spp; so++) {
				uint32_t rng;
				v3 pos, dir;
				primaryRay(j->cam, j->width, j->height, x, y, j->frameIndex, j->frameSeed, so, &rng, &pos, &dir);
				adjointPath(j->s, pos, dir, j->maxDepth, j->rrStart, &rng, delta, gAlb, gEmit);
			}
		}
	}
}

typedef struct { AdjointJob *job; double *gAlb, *gEmit; } AdjointWorker;

static void *adjointWorker(void *arg) {
	AdjointWorker *w = (AdjointWorker *) arg;
	adjointRows(w->job, w->gAlb, w->gEmit);
	return NULL;
}

static int primitiveTotal(const OrScene *s) {
	return s->sphereCount + s->boxCount + s->affineSphereCount + s->affineBoxCount + s->planeCount;
}

int or_adjoint(const OrScene *s, const float *cam, int width, int height, int maxDepth, int rrStart, int frameIndex,
		uint64_t frameSeed, int spp, const float *gradImage, double *gAlb, double *gEmit, int threads) {
	if (!s || !cam || !gradImage || !gAlb || !gEmit || spp < 1 || maxDepth > 32) return -1;
	AdjointJob j = {s, cam, width, height, maxDepth, rrStart, frameIndex, spp, frameSeed, gradImage, gAlb, gEmit, 0};
	pthread_mutex_init(&j.lock, NULL);
	if (threads < 1) threads = 1;
	if (threads > 64) threads = 64;
	size_t n = 3 * (size_t) primitiveTotal(s);
	pthread_t tid[64];
	AdjointWorker w[64];
	for (int i = 1; i < threads; i++) {
		w[i].job = &j;
		w[i].gAlb = (double *) calloc(n ? n : 1, sizeof(double));
		w[i].gEmit = (double *) calloc(n ? n : 1, sizeof(double));
		pthread_create(&tid[i], NULL, adjointWorker, &w[i]);
	}
	adjointRows(&j, gAlb, gEmit);
	for (int i = 1; i < threads; i++) {
		pthread_join(tid[i], NULL);
		for (size_t k = 0; k < n; k++) { gAlb[k] += w[i].gAlb[k]; gEmit[k] += w[i].gEmit[k]; }
		free(w[i].gAlb);
		free(w[i].gEmit);
	}
	pthread_mutex_destroy(&j.lock);
	return 0;
}

/* The frame's paths traced exactly as tracePath (FP32 geometry, the scene's
 * own materials for every sampling decision) with throughput and radiance
 * re-evaluated in double from override materials mat64 (13 doubles each).
 * For Lambertian materials without roulette the path geometry does not depend
 * on albedo or emission, so this image is the frame's estimator as an exact
 * polynomial in those parameters: central differences of it are limited by
 * double, not float, rounding. */
int or_render_fixed_paths64(const OrScene *s, const float *cam, int width, int height, int maxDepth, int frameIndex,
		uint64_t frameSeed, int spp, const double *mat64, double *out) {
	if (!s || !cam || !mat64 || !out || spp < 1) return -1;
	for (int y = 0; y < height; y++)
		for (int x = 0; x < width; x++) {
			uint32_t pixel = (uint32_t) y * (uint32_t) width + (uint32_t) x;
			double acc[3] = {0.0, 0.0, 0.0};
			for (int so = 0; so < spp; so++) {
				uint32_t rng;
				v3 o, d;
				primaryRay(cam, width, height, x, y, frameIndex, frameSeed, so, &rng, &o, &d);
				double T[3] = {1.0, 1.0, 1.0};
				for (int depth = 0; depth < maxDepth; ++depth) {
					Hit h;
					if (!traceClosest(s, o, d, &h, NULL)) {
						for (int c = 0; c < 3; c++) acc[c] += T[c] * (double) s->bg[c];
						break;
					}
					const double *m = h.matIndex >= 0 ? mat64 + 13 * h.matIndex : NULL;
					for (int c = 0; c < 3; c++) acc[c] += m ? T[c] * m[9 + c] : 0.0;
					v3 nrm = h.normal, in = vneg(d);
					float wD = luminance(h.m.diffuse), wR = luminance(h.m.reflective), wT = luminance(h.m.refractive);
					float sum = wD + wR + wT;
					if (sum <= 0.0f) break;
					float xi = rand01(&rng) * sum;
					v3 nd;
					float p;
					int lobe;
					if (xi < wD) { p = wD / sum; nd = vadd(sampleUnitSphere(&rng), nrm); lobe = 0; }
					else if (xi < wD + wR) { p = wR / sum; nd = reflect_dir(nrm, in); lobe = 1; }
					else { p = wT / sum; v3 rd; refract_dir(in, nrm, h.m.ior, &rd); nd = rd; lobe = 2; }
					for (int c = 0; c < 3; c++) T[c] *= (m ? m[3 * lobe + c] : 0.0) / (double) p;
					o = vadd(h.pos, vmuls(nd, EPSILON));
					d = nd;
				}
			}
			for (int c = 0; c < 3; c++) out[3 * pixel + c] = acc[c] / spp;
		}
	return 0;
}

/* The vertices of the frame's paths as tracePath follows them without
 * roulette (FP32 geometry, the scene's own materials for every sampling
 * decision), for re-evaluating the frame's estimator in double under other
 * material parameters: per path (index pixel * spp + sample) the material
 * index, lobe (0 diffuse, 1 reflective, 2 refractive) and lobe probability of
 * each hit vertex, the number of vertices, and how the path ended (1 = its
 * last ray missed and gathered the background, 2 = absorbed at its last
 * vertex (all lobe weights zero), 0 = ran out of depth).  Arrays hold
 * maxDepth vertices per path.  Rows are split over `threads` threads. */
typedef struct {
	const OrScene *s; const float *cam; int width, height, maxDepth, frameIndex, spp; uint64_t frameSeed;
	int32_t *mat; uint8_t *lobe; float *prob; uint8_t *count; uint8_t *end;
	int nextRow; pthread_mutex_t lock;
} RecordJob;

static void recordPath(const OrScene *s, v3 o, v3 d, int maxDepth, uint32_t *rng, int32_t *mat, uint8_t *lobe,
		float *prob, uint8_t *count, uint8_t *end) {
	int n = 0;
	*end = 0;
	for (int depth = 0; depth < maxDepth; ++depth) {
		Hit h;
		if (!traceClosest(s, o, d, &h, NULL)) { *end = 1; break; }
		mat[n] = h.matIndex;
		v3 nrm = h.normal, in = vneg(d);
		float wD = luminance(h.m.diffuse), wR = luminance(h.m.reflective), wT = luminance(h.m.refractive);
		float sum = wD + wR + wT;
		if (sum <= 0.0f) { lobe[n] = 0; prob[n] = 0.0f; n++; *end = 2; break; }
		float xi = rand01(rng) * sum;
		v3 nd;
		if (xi < wD) { prob[n] = wD / sum; nd = vadd(sampleUnitSphere(rng), nrm); lobe[n] = 0; }
		else if (xi < wD + wR) { prob[n] = wR / sum; nd = reflect_dir(nrm, in); lobe[n] = 1; }
		else { prob[n] = wT / sum; v3 rd; refract_dir(in, nrm, h.m.ior, &rd); nd = rd; lobe[n] = 2; }
		n++;
		o = vadd(h.pos, vmuls(nd, EPSILON));
		d = nd;
	}
	*count = (uint8_t) n;
}

static void *recordWorker(void *arg) {
	RecordJob *j = (RecordJob *) arg;
	for (;;) {
		pthread_mutex_lock(&j->lock);
		int y = j->nextRow++;
		pthread_mutex_unlock(&j->lock);
		if (y >= j->height) break;
		for (int x = 0; x < j->width; x++)
			for (int so = 0; so < j->spp; so++) {
				size_t path = ((size_t) y * j->width + x) * j->spp + so;
				size_t v = path * j->maxDepth;
				uint32_t rng;
				v3 o, d;
				primaryRay(j->cam, j->width, j->height, x, y, j->frameIndex, j->frameSeed, so, &rng, &o, &d);
				recordPath(j->s, o, d, j->maxDepth, &rng, j->mat + v, j->lobe + v, j->prob + v, j->count + path,
				           j->end + path);
			}
	}
	return NULL;
}

int or_record_paths(const OrScene *s, const float *cam, int width, int height, int maxDepth, int frameIndex,
		uint64_t frameSeed, int spp, int threads, int32_t *mat, uint8_t *lobe, float *prob, uint8_t *count,
		uint8_t *end) {
	if (!s || !cam || !mat || !lobe || !prob || !count || !end || spp < 1 || maxDepth < 0 || maxDepth > 32) return -1;
	RecordJob j = {s, cam, width, height, maxDepth, frameIndex, spp, frameSeed, mat, lobe, prob, count, end, 0};
	pthread_mutex_init(&j.lock, NULL);
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	pthread_t tid[256];
	for (int i = 1; i < threads; i++) pthread_create(&tid[i], NULL, recordWorker, &j);
	recordWorker(&j);
	for (int i = 1; i < threads; i++) pthread_join(tid[i], NULL);
	pthread_mutex_destroy(&j.lock);
	return 0;
}

/* The per-thread accumulator of the library's calibration kernel (costKernel
 * over its traversal-time operations), restated with this oracle's
 * intersectors: thread tid runs `iterations` iterations from iteration
 * `offset`; each picks a ray of `profile` and `copies` records by hash
 * (PrimitiveCostBenchmark.java:820-841) and folds one operation's result into
 * the float accumulator.  Operation 0 setup, 1 sphere, 2 box, 3 affine sphere,
 * 4 affine box (the reference's routines, GpuKernelSources.java:338-706, with
 * the running best as the accept bound), 5 one node entry distance, 6 a node
 * pair's two entry distances ordered near first (bvhEntryDistance, :736-770;
 * ordering as traceBvh :915-922), 7 plane.  The workload is the calibration
 * layout: profiles x records rays, then spheres, boxes, affine records,
 * planes, node AABBs, node pairs, materials. */
typedef struct { float t; uint32_t rank; int kind, index; } CostBest;

static float costConsume(const CostBest *b) {
	return b->t * 0.000001f + (float) b->kind * 0.000073f + (float) b->rank * 0.0000001f + (float) b->index * 0.0000003f;
}

float or_cost_accumulator(const float *w, int records, int profiles, int op, int profile, int offset, int iterations,
		int copies, int tid) {
	const size_t n = (size_t) records;
	const float *rays = w, *sph = rays + (size_t) profiles * n * 6, *box = sph + n * 5, *aff = box + n * 7;
	const float *pla = aff + n * 22, *aabb = pla + n * 26, *pair = aabb + n * 6, *mats = pair + n * 12;
	const uint32_t mask = (uint32_t) records - 1u;
	float acc = 0.0f;
	for (int i = 0; i < iterations; i++) {
		const uint32_t sample = or_hash_u32((uint32_t) (offset + i) * 0x9E3779B9u + (uint32_t) tid);
		const float *r = rays + ((size_t) profile * n + (sample & mask)) * 6;
		const v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
		for (int c = 0; c < copies; c++) {
			const uint32_t rs = or_hash_u32(sample ^ ((uint32_t) (c + 1) * 0x85EBCA6Bu));
			const int rec = (int) (rs & mask);
			CostBest b = {64.0f + fabsf(o.x) + fabsf(o.y) + fabsf(o.z), (uint32_t) rec, 0, rec};
			Hit h;
			memset(&h, 0, sizeof h);
			h.t = b.t;
			int found = 0, kind = op;
			switch (op) {
				case 1: found = traceSphere(sph + 5 * (size_t) rec, 1, mats, o, d, 0, &h); break;
				case 2: found = traceBox(box + 7 * (size_t) rec, 1, mats, o, d, 0, &h); break;
				case 3: found = traceAffineSphere(aff + 22 * (size_t) rec, 1, mats, o, d, 0, &h); break;
				case 4: found = traceAffineBox(aff + 22 * (size_t) rec, 1, mats, o, d, 0, &h); break;
				case 5: {
					const float e = or_bvh_entry_distance(aabb, rec, o, d, b.t);
					if (isfinite(e)) { found = 1; h.t = e; }
					break;
				}
				case 6: {
					const float *pb = pair + 12 * (size_t) rec;
					const float le = or_bvh_entry_distance(pb, 0, o, d, b.t), re = or_bvh_entry_distance(pb, 1, o, d, b.t);
					const int swap = re < le;
					const int nearN = swap ? 1 : 0, farN = swap ? 0 : 1;
					const float nearE = swap ? re : le, farE = swap ? le : re;
					if (isfinite(nearE)) { found = 1; h.t = nearE; b.index = nearN; }
					else if (isfinite(farE)) { found = 1; h.t = farE; b.index = farN; }
					break;
				}
				case 7: found = tracePlane(pla + 26 * (size_t) rec, 1, mats, o, d, 0, &h); break;
				default: break;
			}
			if (found) { b.t = h.t; b.kind = kind; }
			acc += o.x * 0.00013f + o.y * 0.00017f + o.z * 0.00019f + d.x * 0.00023f + d.y * 0.00029f +
			       d.z * 0.00031f + (float) (rec & 7) * 0.0000001f + costConsume(&b);
		}
	}
	return acc;
}

/* traceReplayKernel body, GpuKernelSources.java:1338-1376, with the per-ray
 * counters it computes kept per ray (perRay: n x 14 u32, may be NULL). */
int or_trace_replay(const OrScene *s, const float *rays, int n, int *flags, float *dist, int *order,
		int *mat, float *normal, uint32_t *perRay) {
	if (!s || !rays || n < 0 || s->stackSize < 1 || s->stackSize > MAX_STACK) return -1;
	for (int r = 0; r < n; r++) {
		const float *ray = rays + 6 * r;
		v3 o = mk(ray[0], ray[1], ray[2]), d = mk(ray[3], ray[4], ray[5]);
		uint64_t st[ST_COUNT] = {0};
		Hit h;
		int f = traceClosest(s, o, d, &h, st);
		st[ST_RAYS] = 1;
		flags[r] = f ? 1 : 0;
		dist[r] = f ? h.t : INFINITY;
		order[r] = f ? h.order : -1;
		mat[r] = f ? h.matIndex : -1;
		normal[3 * r] = f ? h.normal.x : 0.0f;
		normal[3 * r + 1] = f ? h.normal.y : 0.0f;
		normal[3 * r + 2] = f ? h.normal.z : 0.0f;
		if (perRay) for (int i = 0; i < ST_COUNT; i++) perRay[r * ST_COUNT + i] = (uint32_t) st[i];
	}
	return 0;
}
