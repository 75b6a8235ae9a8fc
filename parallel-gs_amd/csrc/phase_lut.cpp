// Phase filter table of super-sampled textures (gs/gs_renderer.cpp:94-313, init_phase_lut).
//
// A super-sampled texture holds one image per sample of the frame buffer it was rendered into.
// Sampling it per sample at a texel phase p (the fractional texel position, relative to the
// grid's mean sample offset) filters the four sample points nearest to p over the 3 x 3 texel
// neighbourhood, with a separable tent of radius f, the largest radius for which no phase has
// a fifth point inside the kernel's support. The table tabulates the four taps (sample id and
// texel offset) and their unorm8 weights for 64 x 64 phases (LINEAR), and the nearest sample
// point for 32 x 32 phases (NEAREST). The reference builds it once per sampling grid on the
// host; so does this backend (Renderer uploads it with the context's planes).
//
// Kept alike with the reference, because the table decides texels: the lattice of sample
// points (the shading kernels' own sample offsets on a 64-step lattice per texel), the tap
// ordering (Chebyshev distance, then vertical distance, sorted with std::sort as the reference
// does, so ties fall the same way with the same standard library), the rule that the first tap
// lies in the centre texel, the float weight arithmetic and the rounding of the weights to
// unorm8 with the remainder on the fourth tap.
#include <algorithm>
#include <climits>
#include <cstdint>
#include "pgs_kernels.h"

namespace
{
constexpr int kLattice = 64; // lattice steps per texel
constexpr int kMaxSamples = 16;

struct Point { int x, y; };
struct Tap
{
	int sample;
	int ox, oy;     // texel offset, -1..1
	int dx, dy;     // |distance| per axis in lattice steps
	int dmax;
};

// Sample points of the grid, in lattice steps inside one texel: the shading kernels' sample
// offsets (1/2^rate_y pixel units) scaled to the lattice; `nearest` moves them to the middle
// of their cell.
int sample_points(uint32_t rate_x, uint32_t rate_y, bool nearest, Point *pts)
{
	const int n = 1 << (rate_x + rate_y);
	for (int i = 0; i < n; i++)
	{
		Point p;
		if (rate_y - rate_x == 2u)
		{
			// Sparse grids: rows of four at x quarters 0, 2, 3, 1, two such columns at 16x.
			static const int kSparse[4] = { 0, 2, 3, 1 };
			p.x = (i / 8) * 4 + kSparse[i % 4];
			p.y = i % 8;
		}
		else
		{
			p.y = (i & 1) + ((i >> 2) & 1) * 2;
			p.x = ((i >> 1) & 1) + ((i >> 3) & 1) * 2;
			if (rate_y - rate_x == 1u) // checkerboards (2x, 8x)
				p.x = p.x * 2 + i % 2;
		}
		const int cell = kLattice >> rate_y;
		p.x *= cell;
		p.y *= cell;
		if (nearest)
		{
			p.x += (cell - 1) >> 1;
			p.y += (cell - 1) >> 1;
		}
		pts[i] = p;
	}
	return n;
}

// Every sample point of the (2 reach + 1)^2 texels around the centre texel, ordered by
// distance to lattice position (x, y).
int taps_at(int x, int y, const Point *pts, int n, int reach, Tap *out)
{
	int k = 0;
	for (int i = 0; i < n; i++)
		for (int oy = -reach; oy <= reach; oy++)
			for (int ox = -reach; ox <= reach; ox++)
			{
				Tap &t = out[k++];
				t.sample = i;
				t.ox = ox;
				t.oy = oy;
				t.dx = std::abs(pts[i].x + kLattice * ox - x);
				t.dy = std::abs(pts[i].y + kLattice * oy - y);
				t.dmax = std::max(t.dx, t.dy);
			}
	std::sort(out, out + k, [](const Tap &a, const Tap &b) {
		return a.dmax == b.dmax ? a.dy < b.dy : a.dmax < b.dmax;
	});
	return k;
}
}

extern "C" void pgs_phase_lut(uint32_t rate_x, uint32_t rate_y, uint32_t *out)
{
	std::fill(out, out + PGS_PHASE_LUT_WORDS, 0u);
	if (rate_y == 0 || rate_x > rate_y || rate_y > 3)
		return;
	Point pts[kMaxSamples], pts_nearest[kMaxSamples];
	const int n = sample_points(rate_x, rate_y, false, pts);
	sample_points(rate_x, rate_y, true, pts_nearest);
	Tap taps[kMaxSamples * 9];

	// Tent radius: the smallest distance of a fifth tap over all phases.
	int falloff = INT_MAX;
	for (int y = 0; y < kLattice; y++)
		for (int x = 0; x < kLattice; x++)
		{
			taps_at(x, y, pts, n, 1, taps);
			falloff = std::min(falloff, taps[4].dmax);
		}

	for (int y = 0; y < kLattice; y++)
		for (int x = 0; x < kLattice; x++)
		{
			const int k = taps_at(x, y, pts, n, 1, taps);
			if (taps[0].ox != 0 || taps[0].oy != 0)
			{
				// The first tap must read the centre texel: the nearest centre-texel point moves to
				// the front, replacing the fourth tap when it was not among the first four.
				for (int i = 1; i < k; i++)
					if (taps[i].ox == 0 && taps[i].oy == 0)
					{
						if (i < 4)
							std::swap(taps[0], taps[i]);
						else
						{
							std::swap(taps[3], taps[i]);
							std::swap(taps[0], taps[3]);
						}
						break;
					}
			}
			uint32_t w0 = 0, w1 = 0;
			float weight[4];
			float sum = 0.0f;
			for (int i = 0; i < 4; i++)
			{
				w0 |= uint32_t(taps[i].sample) << (4 * i);
				w0 |= (uint32_t(taps[i].ox) & 3u) << (4 * i + 16);
				w0 |= (uint32_t(taps[i].oy) & 3u) << (4 * i + 18);
				const float fx = float(taps[i].dx) / float(falloff);
				const float fy = float(taps[i].dy) / float(falloff);
				const float tx = std::max(0.0f, 1.0f - fx);
				const float ty = std::max(0.0f, 1.0f - fy);
				weight[i] = tx * ty;
				sum += weight[i];
			}
			if (sum < 0.001f)
			{
				for (float &w : weight)
					w = 1.0f;
				sum = 4.0f;
			}
			uint32_t used = 0;
			for (int i = 0; i < 3; i++)
			{
				const float scaled = weight[i] * 255.0f;
				uint32_t w = uint32_t(scaled / sum + 0.5f);
				w = std::min(w, 255u - used);
				used += w;
				w1 |= w << (8 * i);
			}
			w1 |= (255u - used) << 24;
			out[2 * (y * kLattice + x)] = w0;
			out[2 * (y * kLattice + x) + 1] = w1;
		}

	// NEAREST: the nearest sample point of the centre texel at the 32 x 32 phases.
	for (int y = 0; y < kLattice / 2; y++)
		for (int x = 0; x < kLattice / 2; x++)
		{
			taps_at(2 * x, 2 * y, pts_nearest, n, 0, taps);
			out[2 * (kLattice * kLattice + y * (kLattice / 2) + x)] = uint32_t(taps[0].sample);
		}
}
