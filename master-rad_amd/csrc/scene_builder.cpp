// Host-side scene packer and BVH builder: the C++ counterpart of the
// reference's GpuScene (src/xyz/marsavic/gfxlab/gpu/GpuScene.java) and
// GpuSceneBuilder (gpu/GpuSceneBuilder.java).  It produces exactly the
// reference's packed wire format (GpuScene.java:24-31, :507-700) so that the
// same arrays can be handed to mrad_upload_scene by any caller.
//
// Builder: full-sweep object SAH over the three centroid axes with the
// reference's cost, tie and leaf rules (GpuScene.java:1414-1499) and the
// primitive-cost weights q = 1 + lambda (c - 1) (BvhBuildConfig.java:67-69).
// Instead of re-sorting every node's primitives per axis, the three axis
// orders are sorted once and stably partitioned at each split.  Because the
// reference's comparator (centroid, then global order) is a total order, a
// node's per-axis order depends only on its primitive set, so the produced
// topology, leaf order and floating-point sums are identical.
//
// Compiled with -ffp-contract=off: every double expression below is evaluated
// in the same order as the Java source it cites.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mrad.h"
#include "mrad_error.hpp"

namespace {

constexpr uint32_t kRefTypeMask = 0xC0000000u;
constexpr uint32_t kRefIndexMask = 0x3FFFFFFFu;
constexpr uint32_t kRefSphere = 0x00000000u;
constexpr uint32_t kRefBox = 0x40000000u;
constexpr uint32_t kRefAffineSphere = 0x80000000u;
constexpr uint32_t kRefAffineBox = 0xC0000000u;
constexpr double kMaxAffineCondition = 65536.0;        // GpuScene.java:19
constexpr double kConditionSlack = 5.0e-4;             // GpuScene.java:22
constexpr double kSplitEpsilon = 1.0e-12;              // GpuScene.java:1415

float minPackedExtent() { return static_cast<float>(std::sqrt(static_cast<double>(std::numeric_limits<float>::min()))); }

// java.lang.Math.min/max for doubles: NaN propagates, -0.0 < +0.0.
inline double jmin(double a, double b) {
	if (a != a) return a;
	if (a == 0.0 && b == 0.0) return std::signbit(a) ? a : b;
	return a <= b ? a : b;
}
inline double jmax(double a, double b) {
	if (a != a) return a;
	if (a == 0.0 && b == 0.0) return std::signbit(a) ? b : a;
	return a >= b ? a : b;
}
// java.lang.Double.compare: total order with -0.0 < +0.0 and NaN last.
inline int jcompare(double a, double b) {
	if (a < b) return -1;
	if (a > b) return 1;
	int64_t ba, bb;
	double ca = (a != a) ? std::numeric_limits<double>::quiet_NaN() : a;
	double cb = (b != b) ? std::numeric_limits<double>::quiet_NaN() : b;
	std::memcpy(&ba, &ca, 8);
	std::memcpy(&bb, &cb, 8);
	return ba == bb ? 0 : (ba < bb ? -1 : 1);
}
inline bool ffinite(double v) { return std::isfinite(static_cast<float>(v)); }
inline float nextDown2(float v) { return std::nextafter(std::nextafter(v, -INFINITY), -INFINITY); }
inline float nextUp2(float v) { return std::nextafter(std::nextafter(v, INFINITY), INFINITY); }

struct V3 { double x, y, z; double get(int a) const { return a == 0 ? x : (a == 1 ? y : z); } };
inline V3 vmin(const V3 &a, const V3 &b) { return {jmin(a.x, b.x), jmin(a.y, b.y), jmin(a.z, b.z)}; }
inline V3 vmax(const V3 &a, const V3 &b) { return {jmax(a.x, b.x), jmax(a.y, b.y), jmax(a.z, b.z)}; }

struct Aabb {
	V3 lo, hi;
	Aabb unite(const Aabb &o) const { return {vmin(lo, o.lo), vmax(hi, o.hi)}; }
	// GpuScene.java:72-75
	double area() const {
		double dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
		return 2.0 * (dx * dy + dx * dz + dy * dz);
	}
};

// Affine (graphics3d/Affine.java), 3x4 row-major.
struct Affine {
	double m[12];
	double a(int r, int c) const { return m[r * 4 + c]; }
	// Affine.java:171-180
	double determinant() const {
		const double m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[4], m11 = m[5], m12 = m[6], m20 = m[8], m21 = m[9], m22 = m[10];
		return (+(m00 * m11 * m22) + (m01 * m12 * m20) + (m02 * m10 * m21) - (m02 * m11 * m20) - (m01 * m10 * m22) - (m00 * m12 * m21));
	}
	// Affine.java:152-160
	Affine inverse() const {
		const double m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3], m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7];
		const double m20 = m[8], m21 = m[9], m22 = m[10], m23 = m[11];
		double det = determinant();
		Affine r;
		r.m[0] = (m11 * m22 - m12 * m21) / det;
		r.m[1] = -(m01 * m22 - m02 * m21) / det;
		r.m[2] = (m01 * m12 - m02 * m11) / det;
		r.m[3] = -(m01 * m12 * m23 + m02 * m13 * m21 + m03 * m11 * m22 - m03 * m12 * m21 - m02 * m11 * m23 - m01 * m13 * m22) / det;
		r.m[4] = -(m10 * m22 - m12 * m20) / det;
		r.m[5] = (m00 * m22 - m02 * m20) / det;
		r.m[6] = -(m00 * m12 - m02 * m10) / det;
		r.m[7] = (m00 * m12 * m23 + m02 * m13 * m20 + m03 * m10 * m22 - m03 * m12 * m20 - m02 * m10 * m23 - m00 * m13 * m22) / det;
		r.m[8] = (m10 * m21 - m11 * m20) / det;
		r.m[9] = -(m00 * m21 - m01 * m20) / det;
		r.m[10] = (m00 * m11 - m01 * m10) / det;
		r.m[11] = -(m00 * m11 * m23 + m01 * m13 * m20 + m03 * m10 * m21 - m03 * m11 * m20 - m01 * m10 * m23 - m00 * m13 * m21) / det;
		return r;
	}
	// Affine.java:162-168
	Affine transposeWithoutTranslation() const {
		Affine r = {{m[0], m[4], m[8], 0.0, m[1], m[5], m[9], 0.0, m[2], m[6], m[10], 0.0}};
		return r;
	}
	// Affine.java:36-42 (at(Vec3.ZERO) evaluated as written)
	V3 atZero() const {
		return {m[0] * 0.0 + m[1] * 0.0 + m[2] * 0.0 + m[3], m[4] * 0.0 + m[5] * 0.0 + m[6] * 0.0 + m[7],
		        m[8] * 0.0 + m[9] * 0.0 + m[10] * 0.0 + m[11]};
	}
};

struct Material { double v[13]; };  // diffuse3, reflective3, refractive3, emittance3, ior

struct MatKey {
	uint64_t b[13];
	bool operator==(const MatKey &o) const { return std::memcmp(b, o.b, sizeof b) == 0; }
};
struct MatKeyHash {
	size_t operator()(const MatKey &k) const {
		uint64_t h = 1469598103934665603ull;
		for (uint64_t w : k.b) { h ^= w; h *= 1099511628211ull; h ^= h >> 29; }
		return static_cast<size_t>(h);
	}
};
// Record equality of MaterialData uses Double.compare per component
// (Color.java:250-258): equal iff same bits, all NaNs equal.
MatKey keyOf(const Material &m) {
	MatKey k;
	for (int i = 0; i < 13; i++) {
		double v = m.v[i];
		if (v != v) v = std::numeric_limits<double>::quiet_NaN();
		std::memcpy(&k.b[i], &v, 8);
	}
	return k;
}

struct SphereRec { V3 c; double r; int mat; };
struct BoxRec { V3 lo, hi; int mat; };
struct PlaneRec { V3 n; double offset; int mat; };
struct AffineRec { Affine inv, invT; V3 center, lo, hi; int mat; };

struct BuildNode { Aabb bounds; int left, right, start, end; };

}  // namespace

struct mrad_scene {
	mrad_bvh_config cfg;
	V3 background{0, 0, 0};
	std::vector<SphereRec> spheres;
	std::vector<BoxRec> boxes;
	std::vector<PlaneRec> planes;
	std::vector<AffineRec> aspheres, aboxes;
	std::vector<Material> materials;
	std::unordered_map<MatKey, int, MatKeyHash> matIndex;
	uint64_t revision = 0;
	bool packed = false;
	// packed wire arrays
	std::vector<float> fMat, fSph, fBox, fPlane, fASph, fABox, fNodeBounds;
	std::vector<int32_t> iNodeData, iRefs, iRoots;
	mrad_bvh_stats stats{};
};

namespace {

bool weighted(const mrad_bvh_config &c) { return c.mode == MRAD_BVH_WEIGHTED_SAH; }
// BvhBuildConfig.java:67-69
double constructionWeight(const mrad_bvh_config &c, double measured) {
	return weighted(c) ? 1.0 + c.lambda * (measured - 1.0) : 1.0;
}

int registerMaterial(mrad_scene *s, const double *m13) {
	Material m;
	std::memcpy(m.v, m13, sizeof m.v);
	MatKey k = keyOf(m);
	auto it = s->matIndex.find(k);
	if (it != s->matIndex.end()) return it->second;
	int idx = static_cast<int>(s->materials.size());
	s->materials.push_back(m);
	s->matIndex.emplace(k, idx);
	return idx;
}

void invalidate(mrad_scene *s) { s->revision++; s->packed = false; }

int fail(const std::string &msg) { return mrad::set_error(MRAD_ERR_INVALID_ARGUMENT, msg); }

bool finite3(const V3 &v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
bool floatRep3(const V3 &v) { return ffinite(v.x) && ffinite(v.y) && ffinite(v.z); }

// GpuScene.java:281-300
bool nondegenerate(const V3 &lo, const V3 &hi) {
	if (!finite3(lo) || !finite3(hi) || !floatRep3(lo) || !floatRep3(hi)) return false;
	const float e = minPackedExtent();
	float lx = (float) lo.x, ly = (float) lo.y, lz = (float) lo.z, hx = (float) hi.x, hy = (float) hi.y, hz = (float) hi.z;
	return lx < hx && ly < hy && lz < hz && (hx - lx) >= e && (hy - ly) >= e && (hz - lz) >= e;
}

// GpuScene.java:407-455
double spectralCondition(const float *mf) {
	double scale = 0.0;
	for (int i = 0; i < 9; i++) scale = jmax(scale, std::fabs((double) mf[i]));
	if (!(scale > 0.0) || !std::isfinite(scale)) return INFINITY;
	double a[9];
	for (int i = 0; i < 9; i++) a[i] = mf[i] / scale;
	double s00 = a[0] * a[0] + a[3] * a[3] + a[6] * a[6];
	double s01 = a[0] * a[1] + a[3] * a[4] + a[6] * a[7];
	double s02 = a[0] * a[2] + a[3] * a[5] + a[6] * a[8];
	double s11 = a[1] * a[1] + a[4] * a[4] + a[7] * a[7];
	double s12 = a[1] * a[2] + a[4] * a[5] + a[7] * a[8];
	double s22 = a[2] * a[2] + a[5] * a[5] + a[8] * a[8];
	double off = s01 * s01 + s02 * s02 + s12 * s12;
	double largest, smallest;
	if (off == 0.0) {
		largest = jmax(s00, jmax(s11, s22));
		smallest = jmin(s00, jmin(s11, s22));
	} else {
		double mean = (s00 + s11 + s22) / 3.0;
		double var = (s00 - mean) * (s00 - mean) + (s11 - mean) * (s11 - mean) + (s22 - mean) * (s22 - mean) + 2.0 * off;
		double p = std::sqrt(var / 6.0);
		double b00 = (s00 - mean) / p, b01 = s01 / p, b02 = s02 / p, b11 = (s11 - mean) / p, b12 = s12 / p, b22 = (s22 - mean) / p;
		double det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02);
		double angle = std::acos(jmax(-1.0, jmin(1.0, det / 2.0))) / 3.0;
		largest = mean + 2.0 * p * std::cos(angle);
		smallest = mean + 2.0 * p * std::cos(angle + 2.0 * M_PI / 3.0);
	}
	if (!(smallest > 0.0) || !std::isfinite(largest)) return INFINITY;
	return std::sqrt(largest / smallest);
}

// GpuScene.java:342-396
int requireAffineMatrices(const Affine &inv, const Affine &invT, const char *label) {
	float il[9], nl[9];
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++) {
			il[r * 3 + c] = (float) inv.a(r, c);
			nl[r * 3 + c] = (float) invT.a(r, c);
		}
	for (int i = 0; i < 9; i++)
		if (!std::isfinite(il[i]) || !std::isfinite(nl[i]))
			return fail(std::string(label) + " matrices must remain finite in the GPU float payload.");
	double det = (double) il[0] * ((double) il[4] * il[8] - (double) il[5] * il[7])
	           - (double) il[1] * ((double) il[3] * il[8] - (double) il[5] * il[6])
	           + (double) il[2] * ((double) il[3] * il[7] - (double) il[4] * il[6]);
	if (det == 0.0 || !std::isfinite(det))
		return fail(std::string(label) + " inverse matrix must be nonsingular after float packing.");
	static const int tr[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};
	for (int i = 0; i < 9; i++)
		if (nl[i] != il[tr[i]])
			return fail(std::string(label) + " normal matrix must be the transpose of its inverse matrix.");
	double cond = spectralCondition(il);
	if (!std::isfinite(cond)) return fail(std::string(label) + " has an invalid condition number");
	if (cond > kMaxAffineCondition * (1.0 + kConditionSlack))
		return fail(std::string(label) + " exceeds the empirically validated float conditioning limit (kappa_2 <= 65536).");
	return MRAD_OK;
}

// GpuScene.java:457-479 (addAffineSphere / addAffineBox with explicit center)
int addAffine(mrad_scene *s, bool box, const Affine &inv, const Affine &invT, V3 center, V3 bmin, V3 bmax,
              const double *mat) {
	const char *label = box ? "Affine box" : "Affine sphere";
	if (!finite3(center) || !floatRep3(center)) return fail(std::string(label) + " world center must be finite in the GPU float payload.");
	V3 pc{(double) (float) center.x, (double) (float) center.y, (double) (float) center.z};
	int rc = requireAffineMatrices(inv, invT, label);
	if (rc != MRAD_OK) return rc;
	V3 lo = vmin(bmin, bmax), hi = vmax(bmin, bmax);
	if (!nondegenerate(lo, hi))
		return fail(std::string(label) + " bounds must retain a supported positive extent on every axis after float packing.");
	if ((float) pc.x < (float) lo.x || (float) pc.x > (float) hi.x || (float) pc.y < (float) lo.y ||
	    (float) pc.y > (float) hi.y || (float) pc.z < (float) lo.z || (float) pc.z > (float) hi.z)
		return fail(std::string(label) + " bounds must enclose the packed world center.");
	AffineRec r{inv, invT, pc, lo, hi, registerMaterial(s, mat)};
	(box ? s->aboxes : s->aspheres).push_back(r);
	invalidate(s);
	return MRAD_OK;
}

// GpuSceneBuilder.java:130-159 followed by sphereBounds/boxBounds (:80-118).
int addAffineFromTransform(mrad_scene *s, bool box, const double *m12, const double *mat) {
	Affine t;
	std::memcpy(t.m, m12, sizeof t.m);
	Affine inv = t.inverse();
	Affine q = {{(double) (float) inv.m[0], (double) (float) inv.m[1], (double) (float) inv.m[2], 0.0,
	             (double) (float) inv.m[4], (double) (float) inv.m[5], (double) (float) inv.m[6], 0.0,
	             (double) (float) inv.m[8], (double) (float) inv.m[9], (double) (float) inv.m[10], 0.0}};
	V3 c0 = t.atZero();
	V3 center{(double) (float) c0.x, (double) (float) c0.y, (double) (float) c0.z};
	Affine invT = q.transposeWithoutTranslation();
	Affine fwd = q.inverse();
	int rc = requireAffineMatrices(q, invT, box ? "Affine box" : "Affine sphere");
	if (rc != MRAD_OK) return rc;
	V3 ext;
	if (box) {
		ext = {std::fabs(fwd.m[0] * 1.0) + std::fabs(fwd.m[1] * 1.0) + std::fabs(fwd.m[2] * 1.0),
		       std::fabs(fwd.m[4] * 1.0) + std::fabs(fwd.m[5] * 1.0) + std::fabs(fwd.m[6] * 1.0),
		       std::fabs(fwd.m[8] * 1.0) + std::fabs(fwd.m[9] * 1.0) + std::fabs(fwd.m[10] * 1.0)};
	} else {
		ext = {std::sqrt(fwd.m[0] * fwd.m[0] + fwd.m[1] * fwd.m[1] + fwd.m[2] * fwd.m[2]),
		       std::sqrt(fwd.m[4] * fwd.m[4] + fwd.m[5] * fwd.m[5] + fwd.m[6] * fwd.m[6]),
		       std::sqrt(fwd.m[8] * fwd.m[8] + fwd.m[9] * fwd.m[9] + fwd.m[10] * fwd.m[10])};
	}
	V3 lo{center.x - ext.x, center.y - ext.y, center.z - ext.z};
	V3 hi{center.x + ext.x, center.y + ext.y, center.z + ext.z};
	return addAffine(s, box, q, invT, center, lo, hi, mat);
}

// ---------------------------------------------------------------- builder

struct Prim { Aabb b; V3 c; double w; double measured; uint32_t ref; int order; };

Aabb conservativeFloat(float lx, float ly, float lz, float hx, float hy, float hz) {
	return {{(double) nextDown2(lx), (double) nextDown2(ly), (double) nextDown2(lz)},
	        {(double) nextUp2(hx), (double) nextUp2(hy), (double) nextUp2(hz)}};
}

class Builder {
public:
	Builder(const mrad_bvh_config &cfg, std::vector<Prim> &prims) : cfg_(cfg), p_(prims) {
		const int n = (int) p_.size();
		for (int a = 0; a < 3; a++) {
			ax_[a].resize(n);
			for (int i = 0; i < n; i++) ax_[a][i] = i;
			std::sort(ax_[a].begin(), ax_[a].end(), [&](int x, int y) {
				int c = jcompare(p_[x].c.get(a), p_[y].c.get(a));
				return c != 0 ? c < 0 : p_[x].order < p_[y].order;
			});
		}
		orig_.resize(n);
		for (int i = 0; i < n; i++) orig_[i] = i;
		mark_.assign(n, 0);
		tmp_.resize(n);
		sufB_.resize(n);
		sufW_.resize(n);
	}

	// Returns the root node; nodes_ holds the tree, leafRefs the leaf order.
	int build() { return buildNode(0, (int) p_.size(), -1, 0); }

	std::vector<BuildNode> nodes_;
	std::vector<int> leafOrder_;   // per node: which axis array holds its order (-1 = original)
	const std::vector<int> &orderArray(int axis) const { return axis < 0 ? orig_ : ax_[axis]; }

private:
	const mrad_bvh_config &cfg_;
	std::vector<Prim> &p_;
	std::vector<int> ax_[3], orig_, tmp_;
	std::vector<char> mark_;
	std::vector<Aabb> sufB_;
	std::vector<double> sufW_;

	int newNode(const Aabb &b, int start, int end, int orderAxis) {
		nodes_.push_back({b, -1, -1, start, end});
		leafOrder_.push_back(orderAxis);
		return (int) nodes_.size() - 1;
	}

	// GpuScene.java:1424-1447
	int buildNode(int start, int end, int orderAxis, int depth) {
		const std::vector<int> &cur = orderArray(orderAxis);
		Aabb bounds = p_[cur[start]].b;
		V3 c0 = p_[cur[start]].c;
		Aabb cb{c0, c0};
		for (int i = start + 1; i < end; i++) {
			const Prim &q = p_[cur[i]];
			bounds = bounds.unite(q.b);
			cb = {vmin(cb.lo, q.c), vmax(cb.hi, q.c)};
		}
		const int count = end - start;
		if (count == 1) return newNode(bounds, start, end, orderAxis);
		int bestAxis = -1, bestLeft = 0;
		double bestCost = 0.0;
		bestSplit(start, end, bounds, cb, &bestAxis, &bestLeft, &bestCost);
		if (bestAxis < 0) return newNode(bounds, start, end, orderAxis);
		double wsum = 0.0;
		for (int i = start; i < end; i++) wsum += p_[cur[i]].w;
		double leafCost = bounds.area() * wsum;
		if (count <= cfg_.leafSize && !(bestCost < leafCost)) return newNode(bounds, start, end, orderAxis);
		partition(start, end, bestAxis, bestLeft);
		const int mid = start + bestLeft;
		int self = newNode(bounds, start, end, orderAxis);
		int l = buildNode(start, mid, bestAxis, depth + 1);
		int r = buildNode(mid, end, bestAxis, depth + 1);
		nodes_[self].left = l;
		nodes_[self].right = r;
		return self;
	}

	// GpuScene.java:1453-1488
	void bestSplit(int start, int end, const Aabb &bounds, const Aabb &cb, int *bestAxis, int *bestLeft, double *bestCost) {
		const int count = end - start;
		const double nodeArea = bounds.area();
		for (int axis = 0; axis < 3; axis++) {
			if (!(cb.hi.get(axis) - cb.lo.get(axis) > kSplitEpsilon)) continue;
			const int *s = ax_[axis].data() + start;
			sufB_[count - 1] = p_[s[count - 1]].b;
			sufW_[count - 1] = p_[s[count - 1]].w;
			for (int i = count - 2; i >= 0; i--) {
				sufB_[i] = sufB_[i + 1].unite(p_[s[i]].b);
				sufW_[i] = sufW_[i + 1] + p_[s[i]].w;
			}
			Aabb pre = p_[s[0]].b;
			double preW = p_[s[0]].w;
			for (int i = 1; i < count; i++) {
				double cost = nodeArea * cfg_.traversalWeight + pre.area() * preW + sufB_[i].area() * sufW_[i];
				if (*bestAxis < 0 || cost < *bestCost) { *bestAxis = axis; *bestLeft = i; *bestCost = cost; }
				pre = pre.unite(p_[s[i]].b);
				preW = preW + p_[s[i]].w;
			}
		}
	}

	// Left set = first `left` entries of the chosen axis order; the other two
	// axis orders are stably partitioned so each keeps its sorted order.
	void partition(int start, int end, int axis, int left) {
		const int *s = ax_[axis].data() + start;
		for (int i = 0; i < end - start; i++) mark_[s[i]] = i < left ? 1 : 0;
		for (int a = 0; a < 3; a++) {
			if (a == axis) continue;
			int *v = ax_[a].data();
			int w = start;
			for (int i = start; i < end; i++) if (mark_[v[i]]) tmp_[w++] = v[i];
			for (int i = start; i < end; i++) if (!mark_[v[i]]) tmp_[w++] = v[i];
			std::memcpy(v + start, tmp_.data() + start, sizeof(int) * (end - start));
		}
	}
};

int leafKind(const std::vector<Prim> &p, const std::vector<int> &order, int start, int end) {
	uint32_t type = p[order[start]].ref & kRefTypeMask;
	for (int i = start + 1; i < end; i++)
		if ((p[order[i]].ref & kRefTypeMask) != type) return -1;
	switch (type) {
		case kRefSphere: return -2;
		case kRefBox: return -3;
		case kRefAffineSphere: return -4;
		default: return -5;
	}
}

// GpuScene.java:810-838
double sahCost(const Builder &b, const std::vector<Prim> &p, int node, bool measured, double tw, const mrad_bvh_config *cfg) {
	const BuildNode &n = b.nodes_[node];
	if (n.left < 0) {
		const std::vector<int> &ord = b.orderArray(b.leafOrder_[node]);
		double c = 0.0;
		for (int i = n.start; i < n.end; i++)
			c += cfg ? constructionWeight(*cfg, p[ord[i]].measured) : (measured ? p[ord[i]].measured : 1.0);
		return c;
	}
	double area = n.bounds.area();
	double l = sahCost(b, p, n.left, measured, tw, cfg), r = sahCost(b, p, n.right, measured, tw, cfg);
	if (!(area > 0.0)) return l + r;
	return tw + b.nodes_[n.left].bounds.area() / area * l + b.nodes_[n.right].bounds.area() / area * r;
}

void writeVec3(std::vector<float> &d, size_t off, const V3 &v) {
	d[off] = (float) v.x; d[off + 1] = (float) v.y; d[off + 2] = (float) v.z;
}

// GpuScene.java:507-601
void packPrimitives(mrad_scene *s) {
	s->fMat.assign(s->materials.size() * 13, 0.0f);
	for (size_t i = 0; i < s->materials.size(); i++)
		for (int k = 0; k < 13; k++) s->fMat[i * 13 + k] = (float) s->materials[i].v[k];
	s->fSph.assign(s->spheres.size() * 5, 0.0f);
	for (size_t i = 0; i < s->spheres.size(); i++) {
		writeVec3(s->fSph, i * 5, s->spheres[i].c);
		s->fSph[i * 5 + 3] = (float) s->spheres[i].r;
		s->fSph[i * 5 + 4] = (float) s->spheres[i].mat;
	}
	s->fBox.assign(s->boxes.size() * 7, 0.0f);
	for (size_t i = 0; i < s->boxes.size(); i++) {
		writeVec3(s->fBox, i * 7, s->boxes[i].lo);
		writeVec3(s->fBox, i * 7 + 3, s->boxes[i].hi);
		s->fBox[i * 7 + 6] = (float) s->boxes[i].mat;
	}
	s->fPlane.assign(s->planes.size() * 26, 0.0f);
	for (size_t i = 0; i < s->planes.size(); i++) {
		float *d = s->fPlane.data() + i * 26;
		d[0] = (float) s->planes[i].n.x; d[1] = (float) s->planes[i].n.y; d[2] = (float) s->planes[i].n.z;
		d[3] = (float) s->planes[i].offset;
		d[18 + 5] = (float) s->planes[i].mat;
		d[18 + 6] = -1.0f;
		d[18 + 7] = -1.0f;
	}
	auto packAffine = [](std::vector<float> &dst, const std::vector<AffineRec> &src) {
		dst.assign(src.size() * 22, 0.0f);
		for (size_t i = 0; i < src.size(); i++) {
			float *d = dst.data() + i * 22;
			const Affine &a = src[i].inv;
			d[0] = (float) a.m[0]; d[1] = (float) a.m[1]; d[2] = (float) a.m[2]; d[3] = (float) src[i].center.x;
			d[4] = (float) a.m[4]; d[5] = (float) a.m[5]; d[6] = (float) a.m[6]; d[7] = (float) src[i].center.y;
			d[8] = (float) a.m[8]; d[9] = (float) a.m[9]; d[10] = (float) a.m[10]; d[11] = (float) src[i].center.z;
			const Affine &t = src[i].invT;
			d[12] = (float) t.m[0]; d[13] = (float) t.m[1]; d[14] = (float) t.m[2];
			d[15] = (float) t.m[4]; d[16] = (float) t.m[5]; d[17] = (float) t.m[6];
			d[18] = (float) t.m[8]; d[19] = (float) t.m[9]; d[20] = (float) t.m[10];
			d[21] = (float) src[i].mat;
		}
	};
	packAffine(s->fASph, s->aspheres);
	packAffine(s->fABox, s->aboxes);
}

// GpuScene.java:614-700 and flattenBvh :772-794
int buildBvh(mrad_scene *s) {
	auto t0 = std::chrono::steady_clock::now();
	const mrad_bvh_config &cfg = s->cfg;
	const size_t count = s->spheres.size() + s->boxes.size() + s->aspheres.size() + s->aboxes.size();
	s->fNodeBounds.clear(); s->iNodeData.clear(); s->iRefs.clear(); s->iRoots.clear();
	std::memset(&s->stats, 0, sizeof s->stats);
	s->stats.leaf_size = cfg.leafSize;
	if (count == 0) return MRAD_OK;
	if (count > kRefIndexMask) return fail("Packed primitive type or index exceeds its assigned bits.");
	std::vector<Prim> prims;
	prims.reserve(count);
	int order = 0;
	for (size_t i = 0; i < s->spheres.size(); i++, order++) {
		const SphereRec &q = s->spheres[i];
		float cx = (float) q.c.x, cy = (float) q.c.y, cz = (float) q.c.z, r = (float) q.r;
		Aabb b = conservativeFloat(cx - r, cy - r, cz - r, cx + r, cy + r, cz + r);
		prims.push_back({b, {}, 0, cfg.sphereWeight, kRefSphere | (uint32_t) i, order});
	}
	for (size_t i = 0; i < s->boxes.size(); i++, order++) {
		const BoxRec &q = s->boxes[i];
		Aabb b = conservativeFloat((float) q.lo.x, (float) q.lo.y, (float) q.lo.z, (float) q.hi.x, (float) q.hi.y, (float) q.hi.z);
		prims.push_back({b, {}, 0, cfg.boxWeight, kRefBox | (uint32_t) i, order});
	}
	for (size_t i = 0; i < s->aspheres.size(); i++, order++) {
		const AffineRec &q = s->aspheres[i];
		Aabb b = conservativeFloat((float) q.lo.x, (float) q.lo.y, (float) q.lo.z, (float) q.hi.x, (float) q.hi.y, (float) q.hi.z);
		prims.push_back({b, {}, 0, cfg.affineSphereWeight, kRefAffineSphere | (uint32_t) i, order});
	}
	for (size_t i = 0; i < s->aboxes.size(); i++, order++) {
		const AffineRec &q = s->aboxes[i];
		Aabb b = conservativeFloat((float) q.lo.x, (float) q.lo.y, (float) q.lo.z, (float) q.hi.x, (float) q.hi.y, (float) q.hi.z);
		prims.push_back({b, {}, 0, cfg.affineBoxWeight, kRefAffineBox | (uint32_t) i, order});
	}
	for (Prim &q : prims) {
		// BvhPrimitive.centroid(): bounds.min().add(bounds.max()).mul(0.5)
		q.c = {(q.b.lo.x + q.b.hi.x) * 0.5, (q.b.lo.y + q.b.hi.y) * 0.5, (q.b.lo.z + q.b.hi.z) * 0.5};
		q.w = constructionWeight(cfg, q.measured);
	}
	Builder b(cfg, prims);
	int root = b.build();
	const int nodeCount = (int) b.nodes_.size();
	s->fNodeBounds.assign((size_t) nodeCount * 6, 0.0f);
	s->iNodeData.assign((size_t) nodeCount * 4, 0);
	s->iRefs.assign(count, 0);
	int nextNode = 0, nextRef = 0, leaves = 0, maxDepth = 0, minOcc = INT32_MAX, maxOcc = 0;
	// Iterative DFS pre-order numbering (left subtree first), same as the recursion.
	struct Item { int node, depth, parentSlot; };
	std::vector<Item> stack;
	stack.push_back({root, 0, -1});
	while (!stack.empty()) {
		Item it = stack.back();
		stack.pop_back();
		const BuildNode &n = b.nodes_[it.node];
		int index = nextNode++;
		if (it.parentSlot >= 0) s->iNodeData[it.parentSlot] = index;
		writeVec3(s->fNodeBounds, (size_t) index * 6, n.bounds.lo);
		writeVec3(s->fNodeBounds, (size_t) index * 6 + 3, n.bounds.hi);
		int32_t *nd = s->iNodeData.data() + (size_t) index * 4;
		maxDepth = std::max(maxDepth, it.depth);
		if (n.left < 0) {
			const std::vector<int> &ord = b.orderArray(b.leafOrder_[it.node]);
			nd[0] = leafKind(prims, ord, n.start, n.end);
			nd[1] = -1;
			nd[2] = nextRef;
			nd[3] = n.end - n.start;
			for (int i = n.start; i < n.end; i++) s->iRefs[nextRef++] = (int32_t) prims[ord[i]].ref;
			leaves++;
			minOcc = std::min(minOcc, n.end - n.start);
			maxOcc = std::max(maxOcc, n.end - n.start);
			continue;
		}
		nd[2] = -1;
		nd[3] = 0;
		// right is pushed first so the left subtree is numbered first
		stack.push_back({n.right, it.depth + 1, index * 4 + 1});
		stack.push_back({n.left, it.depth + 1, index * 4 + 0});
	}
	s->iRoots.assign(1, 0);
	if (nodeCount != 2 * leaves - 1 || nextRef != (int) count)
		return mrad::set_error(MRAD_ERR_ILLEGAL_STATE, "Packed BVH traversal payload is incomplete.");
	mrad_bvh_stats &st = s->stats;
	st.node_count = nodeCount;
	st.root_count = 1;
	st.leaf_count = leaves;
	st.max_depth = maxDepth;
	st.primitive_ref_count = (int) count;
	st.leaf_size = cfg.leafSize;
	st.uniform_sah_cost = sahCost(b, prims, root, false, cfg.traversalWeight, nullptr);
	st.weighted_sah_cost = sahCost(b, prims, root, true, cfg.traversalWeight, nullptr);
	st.generalized_sah_cost = sahCost(b, prims, root, false, cfg.traversalWeight, &cfg);
	st.bytes = (int64_t) s->fNodeBounds.size() * 4 + (int64_t) (s->iNodeData.size() + s->iRefs.size() + s->iRoots.size()) * 4;
	st.min_leaf_occupancy = minOcc == INT32_MAX ? 0 : minOcc;
	st.max_leaf_occupancy = maxOcc;
	st.mean_leaf_occupancy = leaves == 0 ? 0.0 : (double) count / leaves;
	st.build_nanos = (int64_t) std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
	return MRAD_OK;
}

}  // namespace

extern "C" {

int mrad_scene_create(const mrad_bvh_config *cfg, mrad_scene **out) {
	mrad::clear_error();
	if (!cfg || !out) return fail("BVH build config and output handle are required.");
	*out = nullptr;
	mrad_bvh_config c = *cfg;
	if (c.mode != MRAD_BVH_UNIFORM_SAH && c.mode != MRAD_BVH_WEIGHTED_SAH)
		return mrad::set_error(MRAD_ERR_UNSUPPORTED, "Only UNIFORM_SAH and WEIGHTED_SAH BVH modes are built here.");
	if (c.leafSize < 1) return fail("BVH leaf size must be positive.");
	const double ws[5] = {c.traversalWeight, c.sphereWeight, c.boxWeight, c.affineSphereWeight, c.affineBoxWeight};
	for (double w : ws)
		if (!(w > 0.0) || !std::isfinite(w)) return fail("Invalid BVH cost weight.");
	if (c.lambda < 0.0 || !std::isfinite(c.lambda)) return fail("Invalid SAH lambda.");
	// BvhBuildConfig.java:33-43: a weighted mode with lambda zero is uniform.
	if (c.mode == MRAD_BVH_WEIGHTED_SAH && c.lambda == 0.0) c.mode = MRAD_BVH_UNIFORM_SAH;
	else if (c.mode == MRAD_BVH_UNIFORM_SAH && c.lambda != 0.0)
		return fail("An unweighted BVH mode requires the canonical lambda-zero configuration.");
	if (c.mode == MRAD_BVH_WEIGHTED_SAH) {
		const double g[4] = {c.sphereWeight, c.boxWeight, c.affineSphereWeight, c.affineBoxWeight};
		for (double w : g)
			if (!(1.0 + c.lambda * (w - 1.0) > 0.0)) return fail("Invalid generalized primitive weight.");
	}
	mrad_scene *s = new (std::nothrow) mrad_scene();
	if (!s) return mrad::set_error(MRAD_ERR_OUT_OF_MEMORY, "scene allocation failed");
	s->cfg = c;
	*out = s;
	return MRAD_OK;
}

void mrad_scene_destroy(mrad_scene *s) { delete s; }

int mrad_scene_set_background(mrad_scene *s, double r, double g, double b) {
	mrad::clear_error();
	if (!s) return fail("scene is null");
	s->background = {r, g, b};
	invalidate(s);
	return MRAD_OK;
}

// GpuScene.java:205-221
int mrad_scene_add_spheres(mrad_scene *s, int n, const double *center_radius, const double *materials) {
	mrad::clear_error();
	if (!s || n < 0 || (n > 0 && (!center_radius || !materials))) return fail("invalid sphere arrays");
	const float e = minPackedExtent();
	for (int i = 0; i < n; i++) {
		const double *q = center_radius + 4 * i;
		V3 c{q[0], q[1], q[2]};
		double radius = q[3];
		if (!finite3(c) || !floatRep3(c)) return fail("Sphere center must be finite in the GPU float payload (index " + std::to_string(i) + ").");
		float pr = (float) radius;
		if (!(radius > 0.0) || !std::isfinite(radius) || !(pr >= e) || !std::isfinite(pr))
			return fail("Sphere radius must remain finite and at least the minimum packed extent (index " + std::to_string(i) + ").");
		float cx = (float) c.x, cy = (float) c.y, cz = (float) c.z;
		if (!nondegenerate({(double) (cx - pr), (double) (cy - pr), (double) (cz - pr)}, {(double) (cx + pr), (double) (cy + pr), (double) (cz + pr)}))
			return fail("Sphere bounds must retain a supported positive extent on every axis after float packing.");
		s->spheres.push_back({c, radius, registerMaterial(s, materials + 13 * i)});
	}
	invalidate(s);
	return MRAD_OK;
}

// GpuScene.java:223-233
int mrad_scene_add_boxes(mrad_scene *s, int n, const double *min_max, const double *materials) {
	mrad::clear_error();
	if (!s || n < 0 || (n > 0 && (!min_max || !materials))) return fail("invalid box arrays");
	for (int i = 0; i < n; i++) {
		const double *q = min_max + 6 * i;
		V3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]};
		if (!finite3(a) || !finite3(b)) return fail("Box endpoint must be finite.");
		V3 lo = vmin(a, b), hi = vmax(a, b);
		if (!nondegenerate(lo, hi))
			return fail("Box bounds must retain a supported positive extent on every axis after float packing (index " + std::to_string(i) + ").");
		s->boxes.push_back({lo, hi, registerMaterial(s, materials + 13 * i)});
	}
	invalidate(s);
	return MRAD_OK;
}

// GpuScene.java:235-253
int mrad_scene_add_planes(mrad_scene *s, int n, const double *point_normal, const double *materials) {
	mrad::clear_error();
	if (!s || n < 0 || (n > 0 && (!point_normal || !materials))) return fail("invalid plane arrays");
	for (int i = 0; i < n; i++) {
		const double *q = point_normal + 6 * i;
		V3 p{q[0], q[1], q[2]}, nn{q[3], q[4], q[5]};
		if (!finite3(p) || !finite3(nn) || !floatRep3(p) || !floatRep3(nn)) return fail("Plane point and normal must be finite.");
		double l2 = nn.x * nn.x + nn.y * nn.y + nn.z * nn.z;
		if (!(l2 > 0.0) || !std::isfinite(l2)) return fail("Plane normal must be non-zero.");
		double len = std::sqrt(l2);
		V3 u{nn.x / len, nn.y / len, nn.z / len};
		double off = -(u.x * p.x + u.y * p.y + u.z * p.z);
		if (!std::isfinite((float) off)) return fail("Plane offset must remain finite in the GPU float payload.");
		s->planes.push_back({u, off, registerMaterial(s, materials + 13 * i)});
	}
	invalidate(s);
	return MRAD_OK;
}

// GpuScene.java:457-479: explicit inverse, inverse transpose, center and bounds.
int mrad_scene_add_affine(mrad_scene *s, int kind, int n, const double *inverse12, const double *inverse_transpose12,
                          const double *center3, const double *bounds6, const double *materials) {
	mrad::clear_error();
	if (!s || n < 0 || (kind != MRAD_PRIM_AFFINE_SPHERE && kind != MRAD_PRIM_AFFINE_BOX)) return fail("invalid affine arguments");
	if (n > 0 && (!inverse12 || !inverse_transpose12 || !center3 || !bounds6 || !materials)) return fail("invalid affine arrays");
	for (int i = 0; i < n; i++) {
		Affine inv, invT;
		std::memcpy(inv.m, inverse12 + 12 * i, sizeof inv.m);
		std::memcpy(invT.m, inverse_transpose12 + 12 * i, sizeof invT.m);
		const double *c = center3 + 3 * i, *bb = bounds6 + 6 * i;
		int rc = addAffine(s, kind == MRAD_PRIM_AFFINE_BOX, inv, invT, {c[0], c[1], c[2]}, {bb[0], bb[1], bb[2]},
		                   {bb[3], bb[4], bb[5]}, materials + 13 * i);
		if (rc != MRAD_OK) return rc;
	}
	return MRAD_OK;
}

// GpuSceneBuilder.java:37-78 for affine SceneSphere / SceneBox: forward
// transform (3x4 row-major) of the unit sphere or the box [-1,1]^3.
int mrad_scene_add_affine_transforms(mrad_scene *s, int kind, int n, const double *transform12, const double *materials) {
	mrad::clear_error();
	if (!s || n < 0 || (kind != MRAD_PRIM_AFFINE_SPHERE && kind != MRAD_PRIM_AFFINE_BOX)) return fail("invalid affine arguments");
	if (n > 0 && (!transform12 || !materials)) return fail("invalid affine arrays");
	for (int i = 0; i < n; i++) {
		int rc = addAffineFromTransform(s, kind == MRAD_PRIM_AFFINE_BOX, transform12 + 12 * i, materials + 13 * i);
		if (rc != MRAD_OK) return rc;
	}
	return MRAD_OK;
}

int mrad_scene_build(mrad_scene *s) {
	mrad::clear_error();
	if (!s) return fail("scene is null");
	if (s->packed) return MRAD_OK;
	packPrimitives(s);
	int rc = buildBvh(s);
	if (rc != MRAD_OK) return rc;
	s->packed = true;
	return MRAD_OK;
}

int mrad_scene_view(mrad_scene *s, mrad_scene_arrays *out) {
	mrad::clear_error();
	if (!s || !out) return fail("scene and output view are required");
	int rc = mrad_scene_build(s);
	if (rc != MRAD_OK) return rc;
	std::memset(out, 0, sizeof *out);
	out->spheres = s->fSph.data(); out->sphere_count = (int) s->spheres.size();
	out->boxes = s->fBox.data(); out->box_count = (int) s->boxes.size();
	out->planes = s->fPlane.data(); out->plane_count = (int) s->planes.size();
	out->affine_spheres = s->fASph.data(); out->affine_sphere_count = (int) s->aspheres.size();
	out->affine_boxes = s->fABox.data(); out->affine_box_count = (int) s->aboxes.size();
	out->node_bounds = s->fNodeBounds.data(); out->node_count = (int) (s->fNodeBounds.size() / 6);
	out->node_data = s->iNodeData.data();
	out->primitive_refs = s->iRefs.data(); out->primitive_ref_count = (int) s->iRefs.size();
	out->root_indices = s->iRoots.data(); out->root_count = (int) s->iRoots.size();
	out->materials = s->fMat.data(); out->material_count = (int) s->materials.size();
	out->background[0] = (float) s->background.x;
	out->background[1] = (float) s->background.y;
	out->background[2] = (float) s->background.z;
	out->revision = s->revision;
	return MRAD_OK;
}

int mrad_scene_bvh_stats(mrad_scene *s, mrad_bvh_stats *out) {
	mrad::clear_error();
	if (!s || !out) return fail("scene and output stats are required");
	int rc = mrad_scene_build(s);
	if (rc != MRAD_OK) return rc;
	*out = s->stats;
	return MRAD_OK;
}

}  // extern "C"
