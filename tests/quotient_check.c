/* CPU check of the quotient scheme the device slab and sphere tests use:
 * for y = RN(1/b), q = RN(a y), r = a - b q (one fma), RN(q + r y) (one fma)
 * equals the IEEE quotient RN(a/b) whenever |a| and |b| lie in [2^-60, 2^60]
 * (P. Markstein, IBM J. Res. Dev. 34(1), 1990, Theorem 1).
 *
 * usage: quotient_check <random pairs per thread> ; prints "mismatches N". */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float corrected(float a, float b, float y) {
	const float q = a * y;
	return fmaf(fmaf(-q, b, a), y, q);
}

static uint64_t next(uint64_t *s) {
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

/* Random float with exponent in [-60, 60]; a quarter of the significands are
 * adversarial (all ones, zero, near all ones, near zero). */
static float draw(uint64_t *s) {
	const uint64_t z = next(s);
	const int e = -60 + (int) ((z >> 40) % 121u);
	uint32_t m = (uint32_t) z & 0x7FFFFFu;
	switch ((z >> 32) & 7u) {
	case 0: m = 0x7FFFFFu; break;
	case 1: m = 0; break;
	case 2: m = 0x7FFFFFu - ((uint32_t) (z >> 36) & 0xFu); break;
	case 3: m = (uint32_t) (z >> 36) & 0xFu; break;
	default: break;
	}
	return u2f(((uint32_t) (z >> 63) << 31) | ((uint32_t) (e + 127) << 23) | m);
}

typedef struct { uint64_t seed, pairs, bad; } Job;

static void *run(void *arg) {
	Job *j = (Job *) arg;
	uint64_t s = j->seed;
	for (uint64_t i = 0; i < j->pairs; i += 8) {
		const float b = draw(&s), y = 1.0f / b;
		for (int k = 0; k < 8; k++) {
			const float a = draw(&s);
			if (f2u(a / b) != f2u(corrected(a, b, y))) j->bad++;
		}
	}
	return NULL;
}

int main(int argc, char **argv) {
	const uint64_t per = argc > 1 ? strtoull(argv[1], NULL, 10) : 10000000ull;
	enum { kThreads = 4 };
	pthread_t th[kThreads];
	Job jobs[kThreads];
	for (int t = 0; t < kThreads; t++) {
		jobs[t].seed = 0x51u + 977u * (uint64_t) t;
		jobs[t].pairs = per;
		jobs[t].bad = 0;
		pthread_create(&th[t], NULL, run, &jobs[t]);
	}
	uint64_t bad = 0;
	for (int t = 0; t < kThreads; t++) {
		pthread_join(th[t], NULL);
		bad += jobs[t].bad;
	}
	/* every significand of a in [1, 2) against divisors with all-ones significands */
	for (uint32_t k = 0; k < 4; k++) {
		const float b = u2f(0x3F800000u | (0x7FFFFFu - k)), y = 1.0f / b;
		for (uint32_t m = 0; m < 0x800000u; m++) {
			const float a = u2f(0x3F800000u | m);
			if (f2u(a / b) != f2u(corrected(a, b, y))) bad++;
		}
	}
	printf("pairs %llu mismatches %llu\n", (unsigned long long) (per * kThreads + 4ull * 0x800000u),
	       (unsigned long long) bad);
	return bad != 0;
}
