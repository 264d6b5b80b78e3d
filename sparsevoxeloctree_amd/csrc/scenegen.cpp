// Seeded synthetic scenes for the tests and the benchmark (the reference loads OBJ files instead,
// src/Scene.cpp:36-147).  Own RNG (SplitMix64 seeding a PCG32 stream) so that the output does not
// depend on the C++ library's distributions.  Colour: one RGB8 per vertex, a triangle takes the
// colour of its first vertex (the reference colours per material, src/Scene.cpp:164).
//
//   kind 1  UV spheres, 64 longitude x 32 latitude segments = 4032 triangles each (a fan at the
//           north pole, quads elsewhere; the last band's quads touch the south pole and leave one
//           exactly degenerate triangle each), radius U[0.02,0.1], centre U[0,1]^3.
//           param = number of spheres (0 -> 64).
//   kind 2  closed box room [0,1]^3, walls tessellated into (200/d)^2 quads, 400/d randomly rotated
//           and scaled boxes, 200/d spheres of 46x28 segments (2530 triangles), all inside the
//           room.  param = d (0 -> 1).
//   kind 3  param spheres as kind 1 (0 -> 992) plus 4032 per sphere thin slivers with aspect ratio
//           log-uniform in [1,1000], long edge U[0.002,0.02], random orientation, centre U[0,1]^3.
//   kind 4  kind 1 with a triangle budget: param (>= 24) triangles as floor(param / 4032) spheres of
//           64 x 32 segments plus, for the remainder r >= 24, one sphere of 2n x n segments,
//           n = floor(sqrt(r / 4)) (2n(2n-1) <= r triangles).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/svo.h"

namespace {

struct Rng {
	uint64_t state, inc;
	explicit Rng(uint64_t seed) {
		uint64_t s = seed;
		state = mix(s);
		inc = mix(s) | 1ull;
	}
	static uint64_t mix(uint64_t &s) {
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	uint32_t next() { // PCG32 (XSH RR)
		uint64_t old = state;
		state = old * 6364136223846793005ull + inc;
		uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
		uint32_t rot = (uint32_t)(old >> 59u);
		return (xs >> rot) | (xs << ((32u - rot) & 31u));
	}
	double uniform() { return (double)next() * (1.0 / 4294967296.0); } // [0,1)
	double uniform(double a, double b) { return a + (b - a) * uniform(); }
	uint32_t rgb() { return next() & 0x00ffffffu; }
};

struct Out {
	float *xyz;
	uint32_t *rgb;
	uint64_t n = 0;
	void tri(const double *a, const double *b, const double *c, uint32_t col) {
		if (xyz) {
			float *p = xyz + 9 * n;
			for (int k = 0; k < 3; ++k) {
				p[k] = (float)a[k];
				p[3 + k] = (float)b[k];
				p[6 + k] = (float)c[k];
			}
			rgb[n] = col;
		}
		++n;
	}
};

void sphere(Out &o, Rng &rng, const double *c, double r, int nlon, int nlat) {
	const double kPi = 3.14159265358979323846;
	// rings 0..nlat, ring 0 and nlat are the poles (one exact point each)
	std::vector<double> v((size_t)(nlat + 1) * nlon * 3);
	std::vector<uint32_t> col((size_t)(nlat + 1) * nlon);
	for (int k = 0; k <= nlat; ++k)
		for (int m = 0; m < nlon; ++m) {
			double *p = &v[((size_t)k * nlon + m) * 3];
			if (k == 0 || k == nlat) {
				p[0] = c[0];
				p[1] = c[1] + (k == 0 ? r : -r);
				p[2] = c[2];
			} else {
				double th = kPi * k / nlat, ph = 2.0 * kPi * m / nlon;
				p[0] = c[0] + r * std::sin(th) * std::cos(ph);
				p[1] = c[1] + r * std::cos(th);
				p[2] = c[2] + r * std::sin(th) * std::sin(ph);
			}
			col[(size_t)k * nlon + m] = rng.rgb();
		}
	auto V = [&](int k, int m) { return &v[((size_t)k * nlon + (m % nlon)) * 3]; };
	auto C = [&](int k, int m) { return col[(size_t)k * nlon + (m % nlon)]; };
	for (int m = 0; m < nlon; ++m) // north fan
		o.tri(V(0, 0), V(1, m), V(1, m + 1), C(0, 0));
	for (int k = 1; k < nlat; ++k)
		for (int m = 0; m < nlon; ++m) {
			o.tri(V(k, m), V(k + 1, m), V(k + 1, m + 1), C(k, m));
			o.tri(V(k, m), V(k + 1, m + 1), V(k, m + 1), C(k, m));
		}
}

void spheres_kind1(Out &o, Rng &rng, uint64_t n, double rscale = 1.0) {
	for (uint64_t s = 0; s < n; ++s) {
		double r = rng.uniform(0.02, 0.1) * rscale;
		double c[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
		sphere(o, rng, c, r, 64, 32);
	}
}

// Axis-aligned wall at coordinate `at` along `axis`, seg x seg quads.
void wall(Out &o, Rng &rng, int axis, double at, int seg) {
	std::vector<uint32_t> col((size_t)(seg + 1) * (seg + 1));
	for (auto &c : col)
		c = rng.rgb();
	auto P = [&](int a, int b, double *p) {
		double u = (double)a / seg, w = (double)b / seg;
		p[axis] = at;
		p[(axis + 1) % 3] = u;
		p[(axis + 2) % 3] = w;
	};
	for (int a = 0; a < seg; ++a)
		for (int b = 0; b < seg; ++b) {
			double p00[3], p10[3], p11[3], p01[3];
			P(a, b, p00);
			P(a + 1, b, p10);
			P(a + 1, b + 1, p11);
			P(a, b + 1, p01);
			uint32_t c = col[(size_t)a * (seg + 1) + b];
			o.tri(p00, p10, p11, c);
			o.tri(p00, p11, p01, c);
		}
}

void box(Out &o, Rng &rng) {
	double h[3] = {rng.uniform(0.01, 0.05), rng.uniform(0.01, 0.05), rng.uniform(0.01, 0.05)};
	double c[3] = {rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)};
	// uniform random rotation from a normalised quaternion of 4 Gaussian-ish components
	double q[4];
	double n2 = 0.0;
	do {
		n2 = 0.0;
		for (double &x : q) {
			x = rng.uniform(-1.0, 1.0);
			n2 += x * x;
		}
	} while (n2 > 1.0 || n2 < 1e-6);
	double inv = 1.0 / std::sqrt(n2);
	for (double &x : q)
		x *= inv;
	const double w = q[0], x = q[1], y = q[2], z = q[3];
	const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
	                        {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
	                        {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
	double corner[8][3];
	for (int k = 0; k < 8; ++k) {
		double l[3] = {(k & 1) ? h[0] : -h[0], (k & 2) ? h[1] : -h[1], (k & 4) ? h[2] : -h[2]};
		for (int a = 0; a < 3; ++a)
			corner[k][a] = c[a] + R[a][0] * l[0] + R[a][1] * l[1] + R[a][2] * l[2];
	}
	static const int faces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
	for (const auto &f : faces) {
		uint32_t col = rng.rgb();
		o.tri(corner[f[0]], corner[f[1]], corner[f[2]], col);
		o.tri(corner[f[0]], corner[f[2]], corner[f[3]], col);
	}
}

void room_kind2(Out &o, Rng &rng, uint64_t d) {
	if (d == 0)
		d = 1;
	const int seg = (int)std::max<uint64_t>(1, 200 / d);
	for (int axis = 0; axis < 3; ++axis) {
		wall(o, rng, axis, 0.0, seg);
		wall(o, rng, axis, 1.0, seg);
	}
	const uint64_t nbox = std::max<uint64_t>(1, 400 / d), nsph = std::max<uint64_t>(1, 200 / d);
	for (uint64_t b = 0; b < nbox; ++b)
		box(o, rng);
	for (uint64_t s = 0; s < nsph; ++s) {
		double r = rng.uniform(0.02, 0.1);
		double c[3] = {rng.uniform(r, 1.0 - r), rng.uniform(r, 1.0 - r), rng.uniform(r, 1.0 - r)};
		sphere(o, rng, c, r, 46, 28);
	}
}

void slivers(Out &o, Rng &rng, uint64_t n, double lscale = 1.0) {
	for (uint64_t s = 0; s < n; ++s) {
		double len = rng.uniform(0.002, 0.02) * lscale;
		double aspect = std::exp(rng.uniform(0.0, std::log(1000.0)));
		double wid = len / aspect;
		// random unit direction u and a unit vector v perpendicular to it
		double u[3], v[3], t[3];
		double z = rng.uniform(-1.0, 1.0), ph = rng.uniform(0.0, 6.283185307179586);
		double rr = std::sqrt(1.0 - z * z);
		u[0] = rr * std::cos(ph), u[1] = rr * std::sin(ph), u[2] = z;
		t[0] = std::fabs(u[0]) < 0.9 ? 1.0 : 0.0, t[1] = t[0] == 1.0 ? 0.0 : 1.0, t[2] = 0.0;
		v[0] = u[1] * t[2] - u[2] * t[1], v[1] = u[2] * t[0] - u[0] * t[2], v[2] = u[0] * t[1] - u[1] * t[0];
		double vn = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
		double rot = rng.uniform(0.0, 6.283185307179586);
		double w2[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
		for (int a = 0; a < 3; ++a)
			v[a] = (v[a] * vn) * std::cos(rot) + w2[a] * vn * std::sin(rot);
		double c[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
		double a[3], b[3], e[3];
		for (int k = 0; k < 3; ++k) {
			a[k] = c[k] - 0.5 * len * u[k];
			b[k] = c[k] + 0.5 * len * u[k];
			e[k] = c[k] + wid * v[k];
		}
		o.tri(a, b, e, rng.rgb());
	}
}

} // namespace

extern "C" int svo_gen_scene(int kind, uint64_t seed, uint64_t param, float *tri_xyz, uint32_t *tri_rgb, uint64_t *T) {
	if (!T || (tri_xyz && !tri_rgb))
		return SVO_ERR_INVALID_ARGUMENT;
	Rng rng(seed);
	Out o{tri_xyz, tri_rgb};
	switch (kind) {
	case 1:
		spheres_kind1(o, rng, param ? param : 64);
		break;
	case 2:
		room_kind2(o, rng, param);
		break;
	case 3: {
		// param: bits 0..31 sphere count (0 -> 992); bits 32..47 sphere radius scale and bits 48..63
		// sliver length scale, in percent (0 -> 100)
		uint64_t n = param & 0xFFFFFFFFull;
		if (n == 0)
			n = 992;
		uint64_t rs = (param >> 32) & 0xFFFF, ls = (param >> 48) & 0xFFFF;
		spheres_kind1(o, rng, n, rs ? rs / 100.0 : 1.0);
		slivers(o, rng, n * 4032, ls ? ls / 100.0 : 1.0);
		break;
	}
	case 4: {
		const uint64_t full = param / 4032, r = param - full * 4032;
		spheres_kind1(o, rng, full);
		if (r >= 24) {
			const int n = (int)std::floor(std::sqrt((double)r / 4.0));
			double rad = rng.uniform(0.02, 0.1);
			double c[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
			sphere(o, rng, c, rad, 2 * n, n);
		}
		break;
	}
	default:
		return SVO_ERR_INVALID_ARGUMENT;
	}
	*T = o.n;
	return SVO_OK;
}
