// Synthetic GS command-stream generator for the benchmark workloads and the parity tests.
//
// It plays the role of the reference's dump generator (dump/gs_dump_generator.cpp:104-185):
// it emits plain GIF packets (PACKED A+D register writes, PACKED vertex data and IMAGE
// payloads) that any GSInterface implementation consumes byte for byte.
//
// PRNG: xoshiro256** seeded through SplitMix64 (Blackman & Vigna), so the CPU oracle and the
// HIP path read identical bytes for a given seed.
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <cmath>
#include <algorithm>

namespace
{
struct Rng
{
	uint64_t s[4];
	explicit Rng(uint64_t seed)
	{
		for (auto &v : s)
		{
			seed += 0x9e3779b97f4a7c15ull;
			uint64_t z = seed;
			z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
			z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
			v = z ^ (z >> 31);
		}
	}
	static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
	uint64_t next()
	{
		uint64_t r = rotl(s[1] * 5, 7) * 9;
		uint64_t t = s[1] << 17;
		s[2] ^= s[0];
		s[3] ^= s[1];
		s[1] ^= s[2];
		s[0] ^= s[3];
		s[2] ^= t;
		s[3] = rotl(s[3], 45);
		return r;
	}
	// Uniform integer in [lo, hi].
	uint32_t range(uint32_t lo, uint32_t hi) { return lo + uint32_t(next() % (uint64_t(hi) - lo + 1u)); }
	// Uniform float in [0, 1).
	double unit() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Stream
{
	std::vector<uint64_t> q; // pairs of 64-bit words = 128-bit GIF qwords

	void tag(uint32_t nloop, bool eop, bool pre, uint32_t prim, uint32_t flg, uint32_t nreg, uint64_t regs)
	{
		uint64_t lo = uint64_t(nloop & 0x7fff) | (uint64_t(eop) << 15) | (uint64_t(pre) << 46) |
		              (uint64_t(prim & 0x7ff) << 47) | (uint64_t(flg & 3) << 58) | (uint64_t(nreg & 0xf) << 60);
		q.push_back(lo);
		q.push_back(regs);
	}

	// A list of A+D register writes in one PACKED GIFtag.
	void ad(const std::vector<std::pair<uint32_t, uint64_t>> &w)
	{
		tag(uint32_t(w.size()), false, false, 0, 0, 1, 0xe);
		for (auto &r : w)
		{
			q.push_back(r.second);
			q.push_back(r.first);
		}
	}

	void image(const uint8_t *bytes, size_t nbytes)
	{
		// nbytes is a multiple of 16. IMAGE NLOOP is 15 bits, so split long payloads.
		size_t qwords = nbytes / 16;
		size_t off = 0;
		while (off < qwords)
		{
			size_t n = std::min<size_t>(qwords - off, 32767);
			tag(uint32_t(n), false, false, 0, 2, 0, 0);
			size_t base = q.size();
			q.resize(base + n * 2);
			memcpy(q.data() + base, bytes + off * 16, n * 16);
			off += n;
		}
	}
};

enum : uint32_t
{
	CT32 = 0x00, CT24 = 0x01, CT16 = 0x02, CT16S = 0x0a,
	T8 = 0x13, T4 = 0x14, T8H = 0x1b, T4HL = 0x24, T4HH = 0x2c,
	Z32 = 0x30, Z24 = 0x31, Z16 = 0x32, Z16S = 0x3a
};

uint32_t bpp_of(uint32_t psm)
{
	switch (psm)
	{
	case CT24: case Z24: return 24;
	case CT16: case CT16S: case Z16: case Z16S: return 16;
	case T8: case T8H: return 8;
	case T4: case T4HL: case T4HH: return 4;
	default: return 32;
	}
}

uint64_t bitbltbuf(uint32_t sbp, uint32_t sbw, uint32_t spsm, uint32_t dbp, uint32_t dbw, uint32_t dpsm)
{
	return uint64_t(sbp) | uint64_t(sbw) << 16 | uint64_t(spsm) << 24 |
	       uint64_t(dbp) << 32 | uint64_t(dbw) << 48 | uint64_t(dpsm) << 56;
}

uint64_t trxpos(uint32_t ssax, uint32_t ssay, uint32_t dsax, uint32_t dsay)
{
	return uint64_t(ssax) | uint64_t(ssay) << 16 | uint64_t(dsax) << 32 | uint64_t(dsay) << 48;
}

uint64_t trxreg(uint32_t w, uint32_t h) { return uint64_t(w) | uint64_t(h) << 32; }

// PACKED vertex qwords (gs/gs_registers.hpp:443-517 layouts).
void packed_rgbaq(Stream &s, uint32_t rgba)
{
	s.q.push_back(uint64_t(rgba & 0xff) | (uint64_t((rgba >> 8) & 0xff) << 32));
	s.q.push_back(uint64_t((rgba >> 16) & 0xff) | (uint64_t(rgba >> 24) << 32));
}

void packed_xyz2(Stream &s, uint32_t x, uint32_t y, uint32_t z)
{
	s.q.push_back(uint64_t(x & 0xffff) | (uint64_t(y & 0xffff) << 32));
	s.q.push_back(uint64_t(z));
}


uint64_t frame_reg(uint32_t fbp, uint32_t fbw, uint32_t psm, uint32_t fbmsk)
{
	return uint64_t(fbp) | uint64_t(fbw) << 16 | uint64_t(psm) << 24 | uint64_t(fbmsk) << 32;
}

uint64_t zbuf_reg(uint32_t zbp, uint32_t psm, uint32_t zmsk)
{
	return uint64_t(zbp) | uint64_t(psm & 0xf) << 24 | uint64_t(zmsk) << 32;
}

uint64_t test_reg(uint32_t ate, uint32_t atst, uint32_t aref, uint32_t afail, uint32_t zte, uint32_t ztst)
{
	return uint64_t(ate) | uint64_t(atst) << 1 | uint64_t(aref) << 4 | uint64_t(afail) << 12 | uint64_t(zte) << 16 |
	       uint64_t(ztst) << 17;
}

uint64_t scissor_reg(uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1)
{
	return uint64_t(x0) | uint64_t(x1) << 16 | uint64_t(y0) << 32 | uint64_t(y1) << 48;
}

// The last stream generated on each thread (generators may run on several threads at once).
thread_local std::vector<uint8_t> g_last;

const uint8_t *publish(Stream &s, size_t *size)
{
	g_last.assign(s.q.size() * 8, 0);
	memcpy(g_last.data(), s.q.data(), g_last.size());
	*size = g_last.size();
	return g_last.data();
}
}

extern "C"
{
// VRAM transfer microbenchmark: n_h2l host->local uploads cycling through 12 PSMs with
// random w,h in 8..max_dim (multiples of 8), interleaved with one local->local copy after
// every `h2l_per_l2l` uploads (same PSM, page-overlapping source and destination).
// Returns a pointer to the stream (valid until the next generator call on the same thread).
const uint8_t *pgs_gen_transfers(uint64_t seed, uint32_t n_h2l, uint32_t n_l2l, uint32_t max_dim, size_t *size)
{
	static const uint32_t psms[12] = { CT32, CT24, CT16, CT16S, T8, T4, T8H, T4HL, T4HH, Z32, Z24, Z16 };
	Rng rng(seed);
	Stream s;
	std::vector<uint8_t> payload;
	uint32_t h2l_per_l2l = n_l2l ? std::max<uint32_t>(1, n_h2l / n_l2l) : 0;
	uint32_t l2l_done = 0;

	for (uint32_t i = 0; i < n_h2l; i++)
	{
		uint32_t psm = psms[i % 12];
		uint32_t w = rng.range(1, max_dim / 8) * 8;
		uint32_t h = rng.range(1, max_dim / 8) * 8;
		uint32_t dsax = rng.range(0, (2048 - w) / 8) * 8;
		uint32_t dsay = rng.range(0, (2048 - h) / 8) * 8;
		uint32_t min_bw = (dsax + w + 63) / 64;
		if ((psm == T8 || psm == T4) && (min_bw & 1))
			min_bw++;
		uint32_t dbw = std::min<uint32_t>(min_bw + rng.range(0, 3), 62);
		if ((psm == T8 || psm == T4) && (dbw & 1))
			dbw++;
		uint32_t dbp = rng.range(0, 16383);

		s.ad({ { 0x50, bitbltbuf(0, 0, 0, dbp, dbw, psm) },
		       { 0x51, trxpos(0, 0, dsax, dsay) },
		       { 0x52, trxreg(w, h) },
		       { 0x53, 0 } });

		size_t nbytes = size_t(w) * h * bpp_of(psm) / 8;
		payload.resize(nbytes);
		for (size_t b = 0; b < nbytes; b += 8)
		{
			uint64_t r = rng.next();
			memcpy(payload.data() + b, &r, std::min<size_t>(8, nbytes - b));
		}
		s.image(payload.data(), nbytes);

		if (h2l_per_l2l && (i + 1) % h2l_per_l2l == 0 && l2l_done < n_l2l)
		{
			l2l_done++;
			uint32_t lpsm = psms[rng.range(0, 11)];
			uint32_t lw = rng.range(1, max_dim / 16) * 8;
			uint32_t lh = rng.range(1, max_dim / 16) * 8;
			uint32_t ssax = rng.range(8, (2048 - lw - 72) / 8) * 8;
			uint32_t ssay = rng.range(8, (2048 - lh - 72) / 8) * 8;
			uint32_t ldsax = ssax + rng.range(0, 16) * 8 - 64;
			uint32_t ldsay = ssay + rng.range(0, 16) * 8 - 64;
			uint32_t bw = (std::max(ssax, ldsax) + lw + 63) / 64;
			if ((lpsm == T8 || lpsm == T4) && (bw & 1))
				bw++;
			uint32_t bp = rng.range(0, 16383);
			s.ad({ { 0x50, bitbltbuf(bp, bw, lpsm, bp, bw, lpsm) },
			       { 0x51, trxpos(ssax, ssay, ldsax, ldsay) },
			       { 0x52, trxreg(lw, lh) },
			       { 0x53, 2 } });
		}
	}
	return publish(s, size);
}

// Plain rendering workload: one frame of n untextured Gouraud triangles (PRIM=TRIANGLE, IIP=1)
// with vertices uniform over a width x height window at XY offset 2048.0, Z uniform 24-bit,
// RGBA uniform, ZTST=GEQUAL into PSMCT32 at FBP 0 with a PSMZ24 depth buffer.
const uint8_t *pgs_gen_plain(uint64_t seed, uint32_t n_tris, uint32_t width, uint32_t height, size_t *size)
{
	Rng rng(seed);
	Stream s;
	const uint32_t ofs = 2048u * 16u;
	const uint32_t fbw = (width + 63) / 64;
	const uint32_t zbp = (fbw * 64u * height * 4u + 8191u) / 8192u;
	s.ad({ { 0x1a, 1 },                                   // PRMODECONT: AC=1
	       { 0x4c, frame_reg(0, fbw, CT32, 0) },          // FRAME_1
	       { 0x4e, zbuf_reg(zbp, Z24, 0) },               // ZBUF_1
	       { 0x18, uint64_t(ofs) | uint64_t(ofs) << 32 }, // XYOFFSET_1
	       { 0x40, scissor_reg(0, width - 1, 0, height - 1) },
	       { 0x47, test_reg(0, 0, 0, 0, 1, 2) },          // TEST_1: ZTE, GEQUAL
	       { 0x42, 0 },                                   // ALPHA_1
	       { 0x46, 0 },                                   // COLCLAMP
	       { 0x45, 0 } });                                // DTHE
	// PRIM = TRIANGLE | IIP, set through the GIFtag PRE bit; 2 registers per vertex.
	s.tag(n_tris * 3, true, true, 0x3 | 0x8, 0, 2, 0x51);
	for (uint32_t t = 0; t < n_tris * 3; t++)
	{
		uint32_t rgba = uint32_t(rng.next());
		uint32_t x = ofs + rng.range(0, width * 16 - 1);
		uint32_t y = ofs + rng.range(0, height * 16 - 1);
		uint32_t z = uint32_t(rng.next() & 0xffffff);
		packed_rgbaq(s, rgba);
		packed_xyz2(s, x, y, z);
	}
	return publish(s, size);
}

// Textured + blended workload: `frames` frames of n_prims primitives (70% triangles, 30%
// sprites) into a width x height PSMCT32 frame with a PSMZ24 depth buffer. Each frame uploads
// a PSMT8 256x256 texture with uniform indices and a 256-entry PSMCT32 CLUT (CSM1, CLD=1),
// then draws with bilinear MODULATE (TCC=1), blend (Cs - Cd) * As >> 7 + Cd with vertex alpha
// uniform in 0..128, ZTST=GEQUAL with ZMSK=0; 20% of primitives use ST/Q with Q uniform in
// [0.5, 2], the rest UV. Placement: a centre uniform over the window; with pareto_alpha = 0 a
// triangle's vertices lie within +-r of it (r uniform 8..64 px) and a sprite is 8..64 px per
// side; with pareto_alpha > 0 the primitive's area is Pareto(alpha) with a 64 px^2 minimum
// (capped at the window) and the side / radius is its square root.
const uint8_t *pgs_gen_textured(uint64_t seed, uint32_t n_prims, uint32_t frames, uint32_t width, uint32_t height,
                                double pareto_alpha, uint32_t tex_psm, uint32_t blend_preset, size_t *size)
{
	// Blend presets (A, B, C, D): 0 = (Cs, Cd, As, Cd) alpha blend, 1 = (Cs, 0, As, Cd) additive,
	// 2 = (0, Cs, As, Cd) subtractive, 3 = (Cs, Cd, FIX = 64, Cd) fixed 50% blend.
	static const uint64_t presets[4] = {
		uint64_t(0) | 1u << 2 | 0u << 4 | 1u << 6,
		uint64_t(0) | 2u << 2 | 0u << 4 | 1u << 6,
		uint64_t(2) | 0u << 2 | 0u << 4 | 1u << 6,
		uint64_t(0) | 1u << 2 | 2u << 4 | 1u << 6 | uint64_t(64) << 32,
	};
	if (tex_psm != CT32 && tex_psm != T8 && tex_psm != T4)
		tex_psm = T8;
	Rng rng(seed);
	Stream s;
	const uint32_t ofs = 2048u * 16u;
	const uint32_t fbw = (width + 63) / 64;
	const uint32_t zbp = (fbw * 64u * height * 4u + 8191u) / 8192u;
	const uint32_t tbp = (zbp * 2u + 8u) * 32u; // texture after the depth buffer
	const uint32_t cbp = tbp + 32u * 32u + 64u; // CLUT after the texture (at most 256 KiB)
	s.ad({ { 0x1a, 1 },
	       { 0x4c, frame_reg(0, fbw, CT32, 0) },
	       { 0x4e, zbuf_reg(zbp, Z24, 0) },
	       { 0x18, uint64_t(ofs) | uint64_t(ofs) << 32 },
	       { 0x40, scissor_reg(0, width - 1, 0, height - 1) },
	       { 0x47, test_reg(0, 0, 0, 0, 1, 2) },
	       { 0x42, presets[blend_preset & 3u] },
	       { 0x14, uint64_t(1) << 5 | uint64_t(1) << 6 },       // TEX1: MMAG = MMIN = LINEAR
	       { 0x08, 0 },                                          // CLAMP: REPEAT
	       { 0x46, 0 },
	       { 0x45, 0 } });
	const uint32_t tex_bytes = 256u * 256u * bpp_of(tex_psm) / 8u;
	std::vector<uint8_t> payload(tex_bytes);
	const uint32_t clut_w = tex_psm == T4 ? 8u : 16u, clut_h = tex_psm == T4 ? 2u : 16u;
	std::vector<uint8_t> clut(clut_w * clut_h * 4u);
	const double area_cap = double(width) * double(height);
	for (uint32_t f = 0; f < frames; f++)
	{
		for (size_t b = 0; b < payload.size(); b += 8)
		{
			uint64_t r = rng.next();
			memcpy(payload.data() + b, &r, 8);
		}
		s.ad({ { 0x50, bitbltbuf(0, 0, 0, tbp, 4, tex_psm) }, { 0x51, trxpos(0, 0, 0, 0) },
		       { 0x52, trxreg(256, 256) }, { 0x53, 0 } });
		s.image(payload.data(), payload.size());
		for (size_t b = 0; b < clut.size(); b += 8)
		{
			uint64_t r = rng.next();
			memcpy(clut.data() + b, &r, 8);
		}
		if (tex_psm != CT32)
		{
			s.ad({ { 0x50, bitbltbuf(0, 0, 0, cbp, 1, CT32) }, { 0x51, trxpos(0, 0, 0, 0) },
			       { 0x52, trxreg(clut_w, clut_h) }, { 0x53, 0 } });
			s.image(clut.data(), clut.size());
		}
		const uint64_t tex0 = uint64_t(tbp) | uint64_t(4) << 14 | uint64_t(tex_psm) << 20 | uint64_t(8) << 26 |
		                      uint64_t(8) << 30 | uint64_t(1) << 34 | uint64_t(0) << 35 | uint64_t(cbp) << 37 |
		                      uint64_t(CT32) << 51 | uint64_t(tex_psm != CT32 ? 1 : 0) << 61;
		s.ad({ { 0x06, tex0 } });

		uint64_t last_prim = ~0ull;
		std::vector<std::pair<uint32_t, uint64_t>> w;
		for (uint32_t i = 0; i < n_prims; i++)
		{
			const bool tri = rng.range(0, 9) < 7;
			const bool stq = rng.range(0, 4) == 0;
			const uint64_t prim = (tri ? 3u : 6u) | 0x8u | 0x10u | 0x40u | (stq ? 0u : 0x100u);
			if (prim != last_prim)
			{
				w.push_back({ 0x00, prim });
				last_prim = prim;
			}
			const uint32_t cx = rng.range(0, width * 16 - 1), cy = rng.range(0, height * 16 - 1);
			uint32_t extent;
			if (pareto_alpha > 0.0)
			{
				const double u = std::max(rng.unit(), 1e-12);
				const double area = std::min(64.0 * std::pow(u, -1.0 / pareto_alpha), area_cap);
				extent = uint32_t(std::sqrt(area) * 16.0);
			}
			else
				extent = rng.range(8, 64) * 16;
			const uint32_t nv = tri ? 3 : 2;
			const uint32_t q_all = rng.range(0, 1);
			const float q_const = float(0.5 + 1.5 * rng.unit());
			for (uint32_t v = 0; v < nv; v++)
			{
				uint32_t x, y;
				if (tri)
				{
					x = cx + rng.range(0, 2 * extent) - extent;
					y = cy + rng.range(0, 2 * extent) - extent;
				}
				else
				{
					// Sprite: extent wide, 8 px .. extent tall.
					x = v == 0 ? cx : cx + extent;
					y = v == 0 ? cy : cy + rng.range(8 * 16, std::max(8u * 16u, extent));
				}
				const uint32_t rgb = uint32_t(rng.next()) & 0xffffffu;
				const uint32_t a = rng.range(0, 128);
				const uint32_t z = uint32_t(rng.next() & 0xffffff);
				const float q = q_all ? q_const : float(0.5 + 1.5 * rng.unit());
				uint32_t qbits;
				memcpy(&qbits, &q, 4);
				w.push_back({ 0x01, uint64_t(rgb | a << 24) | uint64_t(qbits) << 32 });
				if (stq)
				{
					const float S = float(rng.unit() * 2.0) * q, T = float(rng.unit() * 2.0) * q;
					uint32_t sb, tb;
					memcpy(&sb, &S, 4);
					memcpy(&tb, &T, 4);
					w.push_back({ 0x02, uint64_t(sb) | uint64_t(tb) << 32 });
				}
				else
					w.push_back({ 0x03, uint64_t(rng.range(0, 8191)) | uint64_t(rng.range(0, 8191)) << 16 });
				w.push_back({ 0x05, uint64_t(ofs + x) | uint64_t(ofs + y) << 16 | uint64_t(z) << 32 });
			}
			if (w.size() > 200)
			{
				s.ad(w);
				w.clear();
			}
		}
		if (!w.empty())
			s.ad(w);
	}
	return publish(s, size);
}

// Random uploads with fully arbitrary geometry (any w/h, unaligned origins, wrap-around,
// mixed source/destination PSMs for local copies, strides narrower than the rectangle).
// Used by the parity tests for edge cases.
const uint8_t *pgs_gen_transfers_wild(uint64_t seed, uint32_t n, size_t *size)
{
	static const uint32_t psms[13] = { CT32, CT24, CT16, CT16S, T8, T4, T8H, T4HL, T4HH, Z32, Z24, Z16, Z16S };
	Rng rng(seed);
	Stream s;
	std::vector<uint8_t> payload;
	for (uint32_t i = 0; i < n; i++)
	{
		uint32_t dpsm = psms[rng.range(0, 12)];
		uint32_t w = rng.range(1, 80);
		uint32_t h = rng.range(1, 80);
		uint32_t dsax = rng.range(0, 2047);
		uint32_t dsay = rng.range(0, 2047);
		uint32_t dbw = rng.range(0, 12);
		uint32_t dbp = rng.range(0, 16383);
		bool local = rng.range(0, 3) == 0;
		if (local)
		{
			uint32_t spsm = rng.range(0, 1) ? dpsm : psms[rng.range(0, 12)];
			s.ad({ { 0x50, bitbltbuf(rng.range(0, 16383), rng.range(0, 12), spsm, dbp, dbw, dpsm) },
			       { 0x51, trxpos(rng.range(0, 2047), rng.range(0, 2047), dsax, dsay) },
			       { 0x52, trxreg(w, h) },
			       { 0x53, 2 } });
			continue;
		}

		s.ad({ { 0x50, bitbltbuf(0, 0, 0, dbp, dbw, dpsm) },
		       { 0x51, trxpos(0, 0, dsax, dsay) },
		       { 0x52, trxreg(w, h) },
		       { 0x53, 0 } });
		size_t nbits = size_t(w) * h * bpp_of(dpsm);
		size_t nbytes = ((nbits / 64) * 8 + 15) & ~size_t(15); // payload rounded up to 128 bits
		// Sometimes send the transfer in two IMAGE packets, sometimes leave it short.
		uint32_t mode = rng.range(0, 5);
		if (mode == 0 && nbytes >= 32)
			nbytes -= 16;
		payload.resize(nbytes);
		for (size_t b = 0; b < nbytes; b++)
			payload[b] = uint8_t(rng.next());
		if (mode == 1 && nbytes >= 32)
		{
			size_t half = (nbytes / 32) * 16;
			s.image(payload.data(), half);
			s.image(payload.data() + half, nbytes - half);
		}
		else if (nbytes)
			s.image(payload.data(), nbytes);
	}
	return publish(s, size);
}
}
