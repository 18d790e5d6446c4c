/*
 * FP64 CPU path tracer: restatement of the reference's CPU renderer
 * semantics, used as the converged reference image (relMSE, Monte-Carlo
 * band) and as the CPU baseline timed beside the GPU.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_fp32.c for the usage rule).
 *
 * Follows, in double precision:
 *   graphics3d/raytracers/PathTracer.java:37-71   recursion as a loop: miss ->
 *       background, += throughput * emittance, BSDF sample, afterTime 1e-9,
 *       no origin offset (Ray.pd(p, out));
 *   graphics3d/BSDF.java:52-100                   lobe choice among the
 *       positive-luminance lobes, colour divided by the lobe probability
 *       (a single positive lobe is used as is);
 *   graphics3d/GeometryUtils.java:28-62           reflectedN, refractedNN,
 *       cosine lobe = rejection-sampled unit-ball point normalized + n;
 *   graphics3d/solids/Ball.java:51-64, Box.java:83-131, Solid.java:29-41
 *       (affine solids: ray mapped by the inverse, normal by the inverse
 *       transpose, same t);
 *   playground/HeadlessCpuRenderer.java:62        parallel image rows.
 * Geometry is read from the packed wire arrays (float values held in
 * double); the BVH of the wire arrays is traversed with double slab tests.
 * The reference's RNG (mars-bits Sampler/Hashing) is closed third-party code,
 * so this renderer draws from its own SplitMix64 stream: agreement with the
 * GPU is statistical by construction.  Optional Russian roulette uses the GPU
 * rule (survival clamp(max throughput, 0.05, 0.95)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

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
	int stackSize;
} OrScene;

typedef struct { double x, y, z; } d3;
static inline d3 D(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 add(d3 a, d3 b) { return D(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return D(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 mul(d3 a, double s) { return D(a.x * s, a.y * s, a.z * s); }
static inline d3 hmul(d3 a, d3 b) { return D(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 normalized(d3 v) { double l = sqrt(dot(v, v)); return l > 0 ? mul(v, 1.0 / l) : v; }
static inline double get(d3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
static inline double lum(d3 c) { return c.x * 0.212655 + c.y * 0.715158 + c.z * 0.072187; }

static inline uint64_t splitmix(uint64_t *s) {
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static inline double uni(uint64_t *s) { return (double) (splitmix(s) >> 11) * 0x1p-53; }

typedef struct { double t; int kind, index; d3 n; } Hit64;

/* Ball.firstHit (Ball.java:51-64). */
static int ballHit(d3 c, double r, d3 p, d3 d, double after, double *t) {
	d3 e = sub(c, p);
	double dSqr = dot(d, d);
	double l = dot(e, d) / dSqr;
	double mSqr = l * l - (dot(e, e) - r * r) / dSqr;
	if (mSqr > 0) {
		double m = sqrt(mSqr);
		if (l - m > after) { *t = l - m; return 1; }
		if (l + m > after) { *t = l + m; return 1; }
	}
	return 0;
}

/* Box.firstHit (Box.java:83-131). */
static int boxHit(d3 lo, d3 hi, d3 p, d3 d, double after, double *t, d3 *n) {
	double tEnter = -INFINITY, tExit = INFINITY;
	int ea = -1, xa = -1;
	for (int a = 0; a < 3; a++) {
		double o = get(p, a), dir = get(d, a), mn = get(lo, a), mx = get(hi, a);
		if (dir == 0.0) {
			if (o < mn || o > mx) return 0;
			continue;
		}
		double t0 = (mn - o) / dir, t1 = (mx - o) / dir;
		if (t0 > t1) { double s = t0; t0 = t1; t1 = s; }
		if (t0 > tEnter) { tEnter = t0; ea = a; }
		if (t1 < tExit) { tExit = t1; xa = a; }
		if (tEnter >= tExit) return 0;
	}
	if (ea >= 0 && tEnter > after) {
		double s = get(d, ea) > 0 ? -1.0 : (get(d, ea) < 0 ? 1.0 : 0.0);
		*t = tEnter; *n = D(ea == 0 ? s : 0, ea == 1 ? s : 0, ea == 2 ? s : 0);
		return 1;
	}
	if (xa >= 0 && tExit > after) {
		double s = get(d, xa) > 0 ? 1.0 : (get(d, xa) < 0 ? -1.0 : 0.0);
		*t = tExit; *n = D(xa == 0 ? s : 0, xa == 1 ? s : 0, xa == 2 ? s : 0);
		return 1;
	}
	return 0;
}

/* Affine primitive: local ray = inverse * (p - centre), inverse * d; same t
 * (Solid.transformed, Solid.java:29-41). */
static void affineLocal(const float *a, d3 p, d3 d, d3 *lp, d3 *ld) {
	d3 rel = D(p.x - a[3], p.y - a[7], p.z - a[11]);
	*lp = D(a[0] * rel.x + a[1] * rel.y + a[2] * rel.z, a[4] * rel.x + a[5] * rel.y + a[6] * rel.z,
	        a[8] * rel.x + a[9] * rel.y + a[10] * rel.z);
	*ld = D(a[0] * d.x + a[1] * d.y + a[2] * d.z, a[4] * d.x + a[5] * d.y + a[6] * d.z,
	        a[8] * d.x + a[9] * d.y + a[10] * d.z);
}
static d3 affineNormal(const float *a, d3 ln) {
	const float *m = a + 12;
	return normalized(D(m[0] * ln.x + m[1] * ln.y + m[2] * ln.z, m[3] * ln.x + m[4] * ln.y + m[5] * ln.z,
	                    m[6] * ln.x + m[7] * ln.y + m[8] * ln.z));
}

static void testPrim(const OrScene *s, uint32_t ref, d3 p, d3 d, double after, Hit64 *h) {
	uint32_t type = ref & 0xC0000000u;
	int i = (int) (ref & 0x3FFFFFFFu);
	double t;
	d3 n;
	if (type == 0x00000000u) {
		const float *q = s->spheres + 5 * i;
		d3 c = D(q[0], q[1], q[2]);
		if (ballHit(c, q[3], p, d, after, &t) && t < h->t) {
			h->t = t; h->kind = 0; h->index = i;
			h->n = normalized(sub(add(p, mul(d, t)), c));
		}
	} else if (type == 0x40000000u) {
		const float *q = s->boxes + 7 * i;
		if (boxHit(D(q[0], q[1], q[2]), D(q[3], q[4], q[5]), p, d, after, &t, &n) && t < h->t) {
			h->t = t; h->kind = 1; h->index = i; h->n = n;
		}
	} else {
		int box = type == 0xC0000000u;
		const float *a = (box ? s->affineBoxes : s->affineSpheres) + 22 * i;
		d3 lp, ld;
		affineLocal(a, p, d, &lp, &ld);
		if (box) {
			if (boxHit(D(-1, -1, -1), D(1, 1, 1), lp, ld, after, &t, &n) && t < h->t) {
				h->t = t; h->kind = 3; h->index = i; h->n = affineNormal(a, n);
			}
		} else if (ballHit(D(0, 0, 0), 1.0, lp, ld, after, &t) && t < h->t) {
			h->t = t; h->kind = 2; h->index = i;
			h->n = affineNormal(a, add(lp, mul(ld, t)));
		}
	}
}

static double slabEntry(const float *b, d3 p, d3 d, double after, double tMax) {
	double t0 = after, t1 = tMax;
	for (int a = 0; a < 3; a++) {
		double o = get(p, a), dir = get(d, a);
		double lo = b[a], hi = b[a + 3];
		if (dir == 0.0) {
			if (o < lo || o > hi) return INFINITY;
			continue;
		}
		double ta = (lo - o) / dir, tb = (hi - o) / dir;
		if (ta > tb) { double s = ta; ta = tb; tb = s; }
		tb *= 1.0 + 1e-12;
		if (ta > t0) t0 = ta;
		if (tb < t1) t1 = tb;
		if (t1 < t0) return INFINITY;
	}
	return t0;
}

static int closest(const OrScene *s, d3 p, d3 d, Hit64 *h) {
	const double after = 1e-9;
	h->t = INFINITY; h->kind = -1; h->index = -1;
	if (s->nodeCount > 0) {
		int stack[512]; int size = 0;
		for (int r = 0; r < s->rootCount; r++) stack[size++] = s->roots[r];
		while (size > 0) {
			int node = stack[--size];
			if (!(slabEntry(s->nodeBounds + 6 * node, p, d, after, h->t) < INFINITY)) continue;
			const int *nd = s->nodeData + 4 * node;
			if (nd[0] < 0) {
				for (int k = 0; k < nd[3]; k++) testPrim(s, (uint32_t) s->refs[nd[2] + k], p, d, after, h);
				continue;
			}
			if (size + 2 > 512) { /* degenerate depth: fall back to a linear scan */
				for (int k = 0; k < s->refCount; k++) testPrim(s, (uint32_t) s->refs[k], p, d, after, h);
				size = 0;
				break;
			}
			stack[size++] = nd[1];
			stack[size++] = nd[0];
		}
	} else {
		for (int i = 0; i < s->sphereCount; i++) testPrim(s, (uint32_t) i, p, d, after, h);
		for (int i = 0; i < s->boxCount; i++) testPrim(s, 0x40000000u | (uint32_t) i, p, d, after, h);
		for (int i = 0; i < s->affineSphereCount; i++) testPrim(s, 0x80000000u | (uint32_t) i, p, d, after, h);
		for (int i = 0; i < s->affineBoxCount; i++) testPrim(s, 0xC0000000u | (uint32_t) i, p, d, after, h);
	}
	for (int i = 0; i < s->planeCount; i++) {   /* HalfSpace.firstHit */
		const float *q = s->planes + 26 * i;
		d3 n = D(q[0], q[1], q[2]);
		double den = dot(n, d);
		if (den == 0.0) continue;
		double t = -(dot(n, p) + q[3]) / den;
		if (t > after && t < h->t) { h->t = t; h->kind = 4; h->index = i; h->n = n; }
	}
	return h->kind >= 0;
}

static int materialOf(const OrScene *s, const Hit64 *h) {
	switch (h->kind) {
		case 0: return (int) s->spheres[5 * h->index + 4];
		case 1: return (int) s->boxes[7 * h->index + 6];
		case 2: return (int) s->affineSpheres[22 * h->index + 21];
		case 3: return (int) s->affineBoxes[22 * h->index + 21];
		default: return (int) s->planes[26 * h->index + 23];
	}
}

/* PathTracer.radiance as a loop. */
static d3 radiance64(const OrScene *s, d3 p, d3 d, int maxDepth, int rrStart, uint64_t *rng) {
	d3 T = D(1, 1, 1), L = D(0, 0, 0);
	d3 bg = D(s->bg[0], s->bg[1], s->bg[2]);
	for (int depth = 0; depth < maxDepth; depth++) {
		Hit64 h;
		if (!closest(s, p, d, &h)) { L = add(L, hmul(T, bg)); break; }
		int mi = materialOf(s, &h);
		d3 cD = D(0, 0, 0), cR = cD, cT = cD, Le = cD;
		double ior = 1.5;
		if (mi >= 0) {
			const float *m = s->materials + 13 * mi;
			cD = D(m[0], m[1], m[2]); cR = D(m[3], m[4], m[5]); cT = D(m[6], m[7], m[8]); Le = D(m[9], m[10], m[11]);
			ior = m[12];
		}
		L = add(L, hmul(T, Le));
		double w[3] = {lum(cD), lum(cR), lum(cT)};
		double sum = 0; int cnt = 0;
		for (int k = 0; k < 3; k++) if (w[k] > 0) { sum += w[k]; cnt++; }
		if (cnt == 0) break;
		int lobe = -1; double prob = 1.0;
		if (cnt == 1) {
			for (int k = 0; k < 3; k++) if (w[k] > 0) lobe = k;
		} else {
			double u = uni(rng), acc = 0;
			for (int k = 0; k < 3; k++) {
				if (!(w[k] > 0)) continue;
				acc += w[k] / sum;
				lobe = k; prob = w[k] / sum;
				if (acc > u) break;
			}
		}
		d3 n = h.n, i = mul(d, -1.0), out, col;
		if (lobe == 0) {
			double x, y, z, l2;
			do {
				x = 2 * uni(rng) - 1; y = 2 * uni(rng) - 1; z = 2 * uni(rng) - 1;
				l2 = x * x + y * y + z * z;
			} while (l2 > 1 || l2 == 0);
			double c = 1 / sqrt(l2);
			out = add(D(x * c, y * c, z * c), n);
			col = cD;
		} else if (lobe == 1) {
			out = sub(mul(n, 2 * dot(i, n)), i);
			col = cR;
		} else {
			d3 iu = normalized(i);
			double c1 = dot(iu, n);
			double ri = c1 >= 0 ? ior : -1.0 / ior;
			double c2 = 1 - (1 - c1 * c1) / (ri * ri);
			out = c2 > 0 ? sub(mul(n, c1 - sqrt(c2) * ri), iu) : sub(mul(n, 2 * dot(iu, n)), iu);
			col = cT;
		}
		col = mul(col, 1.0 / prob);
		if (col.x == 0 && col.y == 0 && col.z == 0) break;
		p = add(p, mul(d, h.t));
		d = out;
		T = hmul(T, col);
		if (rrStart >= 0 && depth + 1 >= rrStart && depth + 1 < maxDepth) {
			double m = fmax(T.x, fmax(T.y, T.z));
			double q = fmin(fmax(m, 0.05), 0.95);
			if (uni(rng) >= q) break;
			T = mul(T, 1.0 / q);
		}
	}
	return L;
}

typedef struct {
	const OrScene *s; const float *cam; int width, height, maxDepth, rrStart, spp;
	uint64_t seed; double *mean, *var; int nextRow; pthread_mutex_t lock;
	int x0, y0, cw, ch;     /* the window rendered; mean/var are cw*ch*3, row-major over the window */
} Job64;

static void *worker64(void *arg) {
	Job64 *j = (Job64 *) arg;
	const double aspect = (double) j->width / j->height;
	d3 pos = D(j->cam[0], j->cam[1], j->cam[2]), fwd = D(j->cam[3], j->cam[4], j->cam[5]);
	d3 right = D(j->cam[6], j->cam[7], j->cam[8]), up = D(j->cam[9], j->cam[10], j->cam[11]);
	for (;;) {
		pthread_mutex_lock(&j->lock);
		int y = j->nextRow++;
		pthread_mutex_unlock(&j->lock);
		if (y >= j->ch) break;
		const int wy = y;
		y += j->y0;
		for (int wx = 0; wx < j->cw; wx++) {
			const int x = j->x0 + wx;
			uint64_t rng = j->seed ^ (0xD1B54A32D192ED03ull * (uint64_t) ((uint64_t) y * j->width + x + 1));
			d3 s1 = D(0, 0, 0), s2 = D(0, 0, 0);
			for (int k = 0; k < j->spp; k++) {
				double u = ((x + uni(&rng)) * 2.0 / j->width - 1.0) * aspect;
				double v = 1.0 - (y + uni(&rng)) * 2.0 / j->height;
				d3 dir = add(add(fwd, mul(right, u)), mul(up, v));
				d3 c = radiance64(j->s, pos, dir, j->maxDepth, j->rrStart, &rng);
				s1 = add(s1, c);
				s2 = add(s2, hmul(c, c));
			}
			size_t idx = 3 * ((size_t) wy * j->cw + wx);
			double n = j->spp;
			d3 m = mul(s1, 1.0 / n);
			j->mean[idx] = m.x; j->mean[idx + 1] = m.y; j->mean[idx + 2] = m.z;
			if (j->var) {
				double dn = n > 1 ? n - 1 : 1;
				j->var[idx] = (s2.x - n * m.x * m.x) / dn;
				j->var[idx + 1] = (s2.y - n * m.y * m.y) / dn;
				j->var[idx + 2] = (s2.z - n * m.z * m.z) / dn;
			}
		}
	}
	return NULL;
}

/* Renders spp samples per pixel; mean and (optional) per-pixel sample
 * variance of the single-sample estimator, both W*H*3 doubles. */
int or64_render_window(const OrScene *s, const float *cam, int width, int height, int x0, int y0, int cw, int ch,
		int maxDepth, int rrStart, int spp, uint64_t seed, double *mean, double *var, int threads) {
	if (!s || !cam || !mean || width < 1 || height < 1 || spp < 1) return -1;
	if (x0 < 0 || y0 < 0 || cw < 1 || ch < 1 || x0 + cw > width || y0 + ch > height) return -1;
	Job64 j = {s, cam, width, height, maxDepth, rrStart, spp, seed, mean, var, 0};
	j.x0 = x0; j.y0 = y0; j.cw = cw; j.ch = ch;
	pthread_mutex_init(&j.lock, NULL);
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	pthread_t tid[256];
	for (int i = 1; i < threads; i++) pthread_create(&tid[i], NULL, worker64, &j);
	worker64(&j);
	for (int i = 1; i < threads; i++) pthread_join(tid[i], NULL);
	pthread_mutex_destroy(&j.lock);
	return 0;
}

int or64_render(const OrScene *s, const float *cam, int width, int height, int maxDepth, int rrStart, int spp,
		uint64_t seed, double *mean, double *var, int threads) {
	return or64_render_window(s, cam, width, height, 0, 0, width, height, maxDepth, rrStart, spp, seed, mean, var,
	                          threads);
}

/* Closest hit in FP64 for one ray (distance, kind, index, normal). */
int or64_closest(const OrScene *s, const float *ray6, double *t, int *kind, int *index, double *n3) {
	d3 p = D(ray6[0], ray6[1], ray6[2]), d = D(ray6[3], ray6[4], ray6[5]);
	Hit64 h;
	int f = closest(s, p, d, &h);
	*t = h.t; *kind = h.kind; *index = h.index;
	if (n3) { n3[0] = f ? h.n.x : 0; n3[1] = f ? h.n.y : 0; n3[2] = f ? h.n.z : 0; }
	return f;
}
