// CPU oracle: CRTC scanout (vsync). TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Restates GSInterface::vsync (gs/gs_interface.cpp:4626-4695) and GSRenderer::vsync with its
// helpers (gs/gs_renderer.cpp:4055-5099), and the shaders those passes draw with
// (gs/shaders/sample_circuit.frag, quad.vert + blit_circuit.frag, extwrite.comp, weave.frag),
// one pixel at a time: every fragment pass is a loop over the pixels its scissor or viewport
// covers, in the order the reference records the passes. Float steps round once each, in the
// order the shader source writes them (built with -ffp-contract=off). Where Vulkan leaves a
// result to the implementation this restatement fixes it:
//   - unpackUnorm4x8 gives c / 255 correctly rounded; packUnorm4x8 and writes to UNORM
//     attachments give round(clamp(x, 0, 1) * 255) with ties to even;
//   - linear-clamp sampling uses 8-bit sub-texel weights: texel coordinate s = u * size - 0.5,
//     i0 = floor(s), weight of i0 + 1 = round(256 * frac(s)) / 256, the filtered value being the
//     exact weighted byte sum divided once by 255 * 65536;
//   - fixed-function blending computes src * k + dst * (1 - k) in float;
//   - nearest sampling takes texel floor(u * size), clamped to the image;
//   - dot() sums left to right; smoothstep(e0, e1, x) is t * t * (3 - 2 t);
//   - a field of a different size than the history restarts the deinterlacer's history (the
//     reference would sample fields of different extents together).
#include "oracle_render.hpp"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace orc
{
namespace
{
// The privileged registers used by the scanout: 64-bit registers in 16-byte slots
// (gs/gs_interface.hpp:70-108), fields as in gs/gs_registers.hpp:562-736.
struct Priv
{
	const uint8_t *p;
	uint64_t q(uint32_t slot) const
	{
		uint64_t v;
		memcpy(&v, p + 16u * slot, 8);
		return v;
	}
	static uint32_t f(uint64_t v, uint32_t lo, uint32_t n) { return uint32_t(v >> lo) & ((1u << n) - 1u); }
	// PMODE (slot 0)
	uint32_t EN1() const { return f(q(0), 0, 1); }
	uint32_t EN2() const { return f(q(0), 1, 1); }
	uint32_t MMOD() const { return f(q(0), 5, 1); }
	uint32_t SLBG() const { return f(q(0), 7, 1); }
	uint32_t ALP() const { return f(q(0), 8, 8); }
	// SMODE1 (slot 1), SMODE2 (slot 2)
	uint32_t LC() const { return f(q(1), 3, 7); }
	uint32_t CMOD() const { return f(q(1), 13, 2); }
	uint32_t INT() const { return f(q(2), 0, 1); }
	uint32_t FFMD() const { return f(q(2), 1, 1); }
	// DISPFB1/2 (slots 7, 9), DISPLAY1/2 (slots 8, 10)
	uint32_t FBP(int c) const { return f(q(7 + 2 * c), 0, 9); }
	uint32_t FBW(int c) const { return f(q(7 + 2 * c), 9, 6); }
	uint32_t PSM(int c) const { return f(q(7 + 2 * c), 15, 5); }
	uint32_t DBX(int c) const { return f(q(7 + 2 * c) >> 32, 0, 11); }
	uint32_t DBY(int c) const { return f(q(7 + 2 * c) >> 32, 11, 11); }
	uint32_t DX(int c) const { return f(q(8 + 2 * c), 0, 12); }
	uint32_t DY(int c) const { return f(q(8 + 2 * c), 12, 11); }
	uint32_t MAGH(int c) const { return f(q(8 + 2 * c), 23, 4); }
	uint32_t MAGV(int c) const { return f(q(8 + 2 * c), 27, 2); }
	uint32_t DW(int c) const { return f(q(8 + 2 * c) >> 32, 0, 12); }
	uint32_t DH(int c) const { return f(q(8 + 2 * c) >> 32, 12, 11); }
	// EXTBUF (slot 11), EXTDATA (12), EXTWRITE (13), BGCOLOR (14)
	uint32_t EXBP() const { return f(q(11), 0, 14); }
	uint32_t EXBW() const { return f(q(11), 14, 6); }
	uint32_t FBIN() const { return f(q(11), 20, 2); }
	uint32_t WFFMD() const { return f(q(11), 22, 1); }
	uint32_t EMODA() const { return f(q(11), 23, 2); }
	uint32_t EMODC() const { return f(q(11), 25, 2); }
	uint32_t WDX() const { return f(q(11) >> 32, 0, 11); }
	uint32_t WDY() const { return f(q(11) >> 32, 11, 11); }
	uint32_t SMPH() const { return f(q(12), 23, 4); }
	uint32_t SMPV() const { return f(q(12), 27, 2); }
	uint32_t WW() const { return f(q(12) >> 32, 0, 12); }
	uint32_t WH() const { return f(q(12) >> 32, 12, 11); }
	uint32_t WRITE() const { return f(q(13), 0, 1); }
	uint32_t BG(int ch) const { return f(q(14), 8 * ch, 8); }
};

const uint32_t CMOD_PROGRESSIVE = 0, CMOD_NTSC = 2, CMOD_PAL = 3;

float unpack_unorm(uint32_t c, int ch)
{
	return float((c >> (8 * ch)) & 0xffu) / 255.0f;
}

uint32_t pack_unorm(const float v[4])
{
	uint32_t r = 0;
	for (int ch = 0; ch < 4; ch++)
	{
		float x = std::min(std::max(v[ch], 0.0f), 1.0f) * 255.0f;
		r |= uint32_t(std::nearbyint(x)) << (8 * ch); // default rounding: ties to even
	}
	return r;
}

struct SamplingRect
{
	uint32_t image_w, image_h, valid_w, valid_h;
	uint32_t phase_offset, phase_stride;
};

// gs/gs_renderer.cpp:4114-4172.
SamplingRect compute_circuit_rect(const Priv &pr, uint32_t phase, int c, bool force_progressive)
{
	SamplingRect r{};
	bool is_interlaced = pr.INT() != 0;
	bool alternative_sampling = is_interlaced && !pr.FFMD();
	if (alternative_sampling && force_progressive)
		is_interlaced = false;
	uint32_t DW = pr.DW(c) + 1, DH = pr.DH(c) + 1, MAGH = pr.MAGH(c) + 1, MAGV = pr.MAGV(c) + 1;
	r.image_w = DW / MAGH;
	r.image_h = DH / MAGV;
	r.valid_w = r.image_w;
	if (is_interlaced)
	{
		r.phase_offset = alternative_sampling ? ((pr.DY(c) ^ phase) & 1u) : 0u;
		r.phase_stride = alternative_sampling ? 2u : 1u;
		r.valid_h = (r.image_h - r.phase_offset + 1) >> 1;
		r.image_h = (r.image_h + 1) >> 1;
	}
	else
	{
		r.phase_offset = 0;
		r.phase_stride = 1;
		r.valid_h = r.image_h;
		r.image_h = (r.image_h + 1) & ~1u;
	}
	return r;
}

bool is_16bit(uint32_t psm)
{
	return psm == CT16 || psm == CT16S || psm == Z16 || psm == Z16S;
}

// sample_circuit.frag:50-100: one plane's pixel as RGBA8 and the super-sample valid test.
uint32_t circuit_payload(const uint8_t *vram, uint32_t psm, uint32_t elem, uint32_t slice)
{
	if (is_16bit(psm))
	{
		uint16_t h;
		memcpy(&h, vram + size_t(slice) * VRAM_SIZE + size_t(elem) * 2, 2);
		uint32_t r = (h & 0x1fu) << 3, g = ((h >> 5) & 0x1fu) << 3, b = ((h >> 10) & 0x1fu) << 3;
		uint32_t a = (h >> 15) ? 0xffu : 0x00u; // rgba16_to_rgba32(c, aem 0, ta0 0, ta1 0xff)
		return r | (g << 8) | (b << 16) | (a << 24);
	}
	uint32_t w;
	memcpy(&w, vram + size_t(slice) * VRAM_SIZE + size_t(elem) * 4, 4);
	if (psm != CT32)
		w = (w & ~0xff000000u) | 0x80000000u;
	return w;
}

bool super_sample_is_valid(const uint8_t *vram, uint32_t psm, uint32_t elem)
{
	if (is_16bit(psm))
	{
		uint16_t h;
		memcpy(&h, vram + VRAM_SIZE + size_t(elem) * 2, 2);
		return h == 0xffff;
	}
	uint32_t w;
	memcpy(&w, vram + VRAM_SIZE + size_t(elem) * 4, 4);
	return psm != CT32 ? (w & 0xffffffu) == 0xffffffu : w == ~0u;
}

// sample_crtc_circuit (gs/gs_renderer.cpp:4055-4112) drawing sample_circuit.frag.
std::shared_ptr<ScanImage> sample_circuit(const uint8_t *vram, const Priv &pr, int c, const SamplingRect &rect,
                                          uint32_t w, uint32_t h, uint32_t super_samples)
{
	auto img = std::make_shared<ScanImage>(w, h); // cleared to 0 where not drawn
	uint32_t vw = rect.valid_w, vh = rect.valid_h;
	if (super_samples > 1)
	{
		vw *= 2;
		vh *= 2;
	}
	const uint32_t psm = pr.PSM(c);
	for (uint32_t y = 0; y < std::min(vh, h); y++)
		for (uint32_t x = 0; x < std::min(vw, w); x++)
		{
			uint32_t ssx = x, ssy = y;
			uint32_t sx = super_samples >= 4 ? ssx >> 1 : ssx;
			uint32_t sy = super_samples >= 4 ? ssy >> 1 : ssy;
			uint32_t cx = sx * 1u + pr.DBX(c);
			uint32_t cy = sy * rect.phase_stride + (pr.DBY(c) + rect.phase_offset);
			uint32_t elem = swizzle(cx, cy, pr.FBP(c) * 32u, pr.FBW(c), psm);
			uint32_t color;
			if (super_samples == 1)
				color = circuit_payload(vram, psm, elem, 0);
			else if (super_sample_is_valid(vram, psm, elem))
			{
				uint32_t quad_offset = super_samples != 8 ? (ssy & 1u) + (ssx & 1u) * 2u : (ssx & 1u) + (ssy & 1u) * 2u;
				const uint32_t NUM = super_samples / 4;
				uint32_t base_slice = 2 + NUM * quad_offset;
				float acc[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
				for (uint32_t i = 0; i < NUM; i++)
				{
					uint32_t s = circuit_payload(vram, psm, elem, base_slice + i);
					for (int ch = 0; ch < 4; ch++)
						acc[ch] += unpack_unorm(s, ch);
				}
				for (int ch = 0; ch < 4; ch++)
					acc[ch] /= float(NUM);
				color = pack_unorm(acc);
			}
			else
				color = circuit_payload(vram, psm, elem, 0);
			img->px[size_t(y) * w + x] = color;
		}
	return img;
}

struct Rect
{
	int32_t x, y;
	uint32_t w, h;
};

// Linear-clamp texel fetch of blit_circuit.frag at texture coordinate (u, v).
void sample_linear(const ScanImage &img, float u, float v, float out[4])
{
	auto axis = [](float t, uint32_t size, uint32_t &i0, uint32_t &i1, uint32_t &w8) {
		float s = t * float(size) - 0.5f;
		int32_t fixed = int32_t(std::floor(s * 256.0f + 0.5f));
		int32_t i = fixed >> 8; // floor division by 256
		w8 = uint32_t(fixed - i * 256);
		auto clampi = [&](int32_t k) { return uint32_t(std::min(std::max(k, 0), int32_t(size) - 1)); };
		i0 = clampi(i);
		i1 = clampi(i + 1);
	};
	uint32_t x0, x1, y0, y1, a, b;
	axis(u, img.w, x0, x1, a);
	axis(v, img.h, y0, y1, b);
	for (int ch = 0; ch < 4; ch++)
	{
		auto c = [&](uint32_t x, uint32_t y) { return (img.px[size_t(y) * img.w + x] >> (8 * ch)) & 0xffu; };
		uint32_t sum = (256 - b) * ((256 - a) * c(x0, y0) + a * c(x1, y0)) + b * ((256 - a) * c(x0, y1) + a * c(x1, y1));
		out[ch] = float(sum) / (255.0f * 65536.0f);
	}
}

// One full-viewport draw of quad.vert + blit_circuit.frag into `dst` (gs/gs_renderer.cpp:
// 4806-4900). blend: 0 replace, 1 constant alpha `alp`, 2 source alpha.
void blit(ScanImage &dst, const ScanImage &src, float vx, float vy, float vw, float vh, int blend, float alp)
{
	for (uint32_t py = 0; py < dst.h; py++)
		for (uint32_t px = 0; px < dst.w; px++)
		{
			float cx = float(px) + 0.5f, cy = float(py) + 0.5f;
			if (cx < vx || cx >= vx + vw || cy < vy || cy >= vy + vh)
				continue;
			float u = (cx - vx) / vw, v = (cy - vy) / vh;
			float s[4];
			sample_linear(src, u, v, s);
			uint32_t &d = dst.px[size_t(py) * dst.w + px];
			float o[4];
			if (blend == 0)
				memcpy(o, s, sizeof(o));
			else
			{
				float k = blend == 1 ? alp : s[3];
				for (int ch = 0; ch < 3; ch++)
					o[ch] = s[ch] * k + unpack_unorm(d, ch) * (1.0f - k);
				o[3] = s[3];
			}
			d = pack_unorm(o);
		}
}

// extwrite.comp over `view` sampled nearest-clamp.
void extwrite(uint8_t *vram, bool invalidate_super_sample, const ScanImage &view, uint32_t res_x, uint32_t res_y,
              uint32_t wdx, uint32_t wdy, float uv_base[2], float uv_scale[2], uint32_t exbp, uint32_t exbw,
              uint32_t emoda, uint32_t emodc)
{
	for (uint32_t gy = 0; gy < res_y; gy++)
		for (uint32_t gx = 0; gx < res_x; gx++)
		{
			uint32_t elem = swizzle(gx + wdx, gy + wdy, exbp, exbw, CT32);
			float u = uv_base[0] + float(gx) * uv_scale[0];
			float v = uv_base[1] + float(gy) * uv_scale[1];
			int32_t tx = int32_t(std::floor(u * float(view.w)));
			int32_t ty = int32_t(std::floor(v * float(view.h)));
			tx = std::min(std::max(tx, 0), int32_t(view.w) - 1);
			ty = std::min(std::max(ty, 0), int32_t(view.h) - 1);
			uint32_t in = view.px[size_t(ty) * view.w + tx];
			float r = unpack_unorm(in, 0), g = unpack_unorm(in, 1), b = unpack_unorm(in, 2), a = unpack_unorm(in, 3);

			float Ey = 0.587f * g + 0.114f * b + 0.299f * r;
			float Ecb = -0.311f * g + 0.500f * b - 0.169f * r;
			float Ecr = -0.419f * g - 0.081f * b + 0.500f * r;
			float Y = (219.0f * Ey + 16.0f) / 255.0f;
			float Cb = (224.0f * Ecb + 128.0f) / 255.0f;
			float Cr = (224.0f * Ecr + 128.0f) / 255.0f;

			float color[4];
			switch (emodc)
			{
			case 0: color[0] = r; color[1] = g; color[2] = b; break;
			case 1: color[0] = color[1] = color[2] = Y; break;
			case 2: color[0] = Y; color[1] = Cb; color[2] = Cr; break;
			default: color[0] = color[1] = color[2] = a; break;
			}
			switch (emoda)
			{
			case 0: color[3] = a; break;
			case 1: color[3] = Y; break;
			case 2: color[3] = 0.5f * Y; break;
			default: color[3] = 0.0f; break;
			}
			uint32_t word = pack_unorm(color);
			memcpy(vram + size_t(elem) * 4, &word, 4);
			if (invalidate_super_sample)
			{
				uint32_t zero = 0;
				memcpy(vram + VRAM_SIZE + size_t(elem) * 4, &zero, 4);
			}
		}
}

float luma_abs_diff(uint32_t p, uint32_t q)
{
	float d[3];
	for (int ch = 0; ch < 3; ch++)
		d[ch] = unpack_unorm(p, ch) - unpack_unorm(q, ch);
	return std::fabs(d[0] * 0.299f + d[1] * 0.587f + d[2] * 0.114f);
}

// fastmad_deinterlace + weave.frag (gs/gs_renderer.cpp:5057-5099).
std::shared_ptr<ScanImage> weave(const std::shared_ptr<ScanImage> f[4], uint32_t phase)
{
	const uint32_t W = f[0]->w, H = f[0]->h * 2;
	auto out = std::make_shared<ScanImage>(W, H);
	const int height_minus_1 = int(H) - 1;
	auto at = [](const ScanImage &img, uint32_t x, uint32_t y) { return img.px[size_t(y) * img.w + x]; };
	for (uint32_t y = 0; y < H; y++)
		for (uint32_t x = 0; x < W; x++)
		{
			uint32_t c;
			if ((y & 1u) == phase)
				c = at(*f[0], x, y >> 1);
			else if (y == 0 || int(y) == height_minus_1)
				c = at(*f[1], x, y >> 1);
			else
			{
				uint32_t up = (y - 1) >> 1;
				uint32_t cur0 = at(*f[0], x, up), cur1 = at(*f[0], x, up + 1);
				uint32_t prev0 = at(*f[2], x, up), prev1 = at(*f[2], x, up + 1);
				uint32_t field1 = at(*f[1], x, y >> 1), field3 = at(*f[3], x, y >> 1);
				float Mh = luma_abs_diff(cur0, prev0);
				float Ml = luma_abs_diff(cur1, prev1);
				float Mc = luma_abs_diff(field1, field3);
				float diff = std::max(std::min(Mh, Ml), Mc);
				float t = (diff - 0.04f) / (0.06f - 0.04f);
				t = std::min(std::max(t, 0.0f), 1.0f);
				float bob_factor = t * t * (3.0f - 2.0f * t);
				float o[4];
				for (int ch = 0; ch < 3; ch++)
				{
					float bob = 0.5f * (unpack_unorm(cur0, ch) + unpack_unorm(cur1, ch));
					o[ch] = unpack_unorm(field1, ch) * (1.0f - bob_factor) + bob * bob_factor;
				}
				o[3] = 1.0f;
				c = pack_unorm(o);
			}
			out->px[size_t(y) * W + x] = c;
		}
	return out;
}

void emit(const std::shared_ptr<ScanImage> &img, ScanoutOut &out, std::vector<uint32_t> &image)
{
	out.has_image = img != nullptr;
	out.width = img ? img->w : 0;
	out.height = img ? img->h : 0;
	image = img ? img->px : std::vector<uint32_t>();
}
} // namespace

bool GS::vsync_can_skip(const VsyncInfo &info) const
{
	// gs/gs_interface.cpp:4692-4695, gs/gs_renderer.cpp:4174-4187.
	Priv pr{ priv.data() };
	bool is_interlaced = pr.INT() != 0;
	if (is_interlaced && !pr.FFMD() && info.force_progressive)
		is_interlaced = false;
	return complex_scanmsk_timeout == 0 && !is_interlaced;
}

int GS::vsync(const VsyncInfo &info_in, ScanoutOut &out, std::vector<uint32_t> &image)
{
	out = ScanoutOut{};
	image.clear();
	flush();
	Priv pr{ priv.data() };
	const uint32_t rate_x_log2 = this->rate_x_log2(), rate_y_log2 = this->rate_y_log2();

	// GSInterface::vsync (gs/gs_interface.cpp:4626-4690); no promoted back buffers.
	if (info_in.high_resolution_scanout && rate_y_log2 && pr.FFMD() && pr.INT() && pr.CMOD() != CMOD_PROGRESSIVE)
		return -1000; // field-aware rendering would be switched on
	for (int c = 0; c < 2; c++)
	{
		uint32_t psm = pr.PSM(c);
		if ((c == 0 ? pr.EN1() : pr.EN2()) && psm != CT32 && psm != CT24 && psm != CT16 && psm != CT16S)
			return -1000; // only colour formats are displayed
	}
	VsyncInfo info = info_in;
	if (complex_scanmsk_timeout)
	{
		info.force_progressive = 0;
		complex_scanmsk_timeout--;
	}

	// GSRenderer::vsync (gs/gs_renderer.cpp:4189-5055).
	const bool force_progressive = info.force_progressive != 0;
	const bool overscan = info.overscan != 0;
	const bool anti_blur = info.anti_blur != 0;
	bool high_resolution_scanout = info.high_resolution_scanout != 0;
	bool is_interlaced = pr.INT() != 0;
	if (is_interlaced && !pr.FFMD() && force_progressive)
		is_interlaced = false;
	bool force_deinterlace = !high_resolution_scanout && (pr.FFMD() && pr.INT() && pr.CMOD() != CMOD_PROGRESSIVE);
	bool alternative_sampling = pr.INT() && !pr.FFMD();
	bool double_strike = !pr.INT() && pr.CMOD() != CMOD_PROGRESSIVE;
	if (!force_progressive || pr.WRITE() || !rate_x_log2 || !rate_y_log2 || force_deinterlace)
		high_resolution_scanout = false;
	bool field_aware_rendering =
	    high_resolution_scanout && rate_y_log2 && pr.FFMD() && pr.INT() && pr.CMOD() != CMOD_PROGRESSIVE;
	uint32_t super_samples = high_resolution_scanout ? 1u << (rate_x_log2 + rate_y_log2) : 1u;

	uint32_t phase = 0, clock_divider = 1, mode_width = 0, mode_height = 0;
	int scan_offset_x = 0, scan_offset_y = 0;
	const bool doubled_lines = !is_interlaced && !double_strike && !force_deinterlace;
	if ((pr.CMOD() == CMOD_NTSC || (pr.CMOD() == CMOD_PROGRESSIVE && !pr.INT())) && pr.LC() == 32)
	{
		mode_width = overscan ? 712 : 640;
		mode_height = overscan ? 240 : 224;
		scan_offset_x = overscan ? 123 : 159;
		scan_offset_y = overscan ? 17 : 25;
		if (doubled_lines)
		{
			mode_height *= 2;
			scan_offset_y *= 2;
		}
		clock_divider = pr.CMOD() == CMOD_PROGRESSIVE ? 2 : 4;
	}
	else if (pr.CMOD() == CMOD_PAL && pr.LC() == 32)
	{
		mode_width = overscan ? 712 : 640;
		mode_height = overscan ? 288 : 256;
		scan_offset_x = overscan ? 133 : 169;
		scan_offset_y = overscan ? 21 : 36;
		if (doubled_lines)
		{
			mode_height *= 2;
			scan_offset_y *= 2;
		}
		clock_divider = 4;
	}
	else if (pr.CMOD() == CMOD_PROGRESSIVE && pr.LC() == 22)
	{
		mode_width = pr.INT() ? 1920 : 1280;
		mode_height = pr.INT() ? 540 : 720;
		scan_offset_x = pr.INT() ? 236 : 300;
		scan_offset_y = 0;
		if (doubled_lines)
		{
			mode_height *= 2;
			scan_offset_y *= 2;
		}
		clock_divider = 1;
		high_resolution_scanout = false;
		field_aware_rendering = false;
		super_samples = 1;
	}
	else if (pr.CMOD() == CMOD_PROGRESSIVE && pr.LC() == 15)
	{
		mode_width = 640;
		mode_height = 480;
		scan_offset_x = 138;
		scan_offset_y = 34;
		clock_divider = 2;
	}
	else if (pr.LC() == 29)
	{
		mode_width = 640;
		mode_height = 480;
		scan_offset_x = 318;
		scan_offset_y = 50;
		clock_divider = 4;
	}
	else
		return 0; // unknown video format: empty result

	if (alternative_sampling && !force_progressive)
		phase = info.phase;

	bool EN1 = pr.EN1(), EN2 = pr.EN2();
	uint32_t MMOD = pr.MMOD(), ALP = pr.ALP(), SLBG = pr.SLBG();
	int32_t crtc_shift_x = INT32_MAX, crtc_shift_y = INT32_MAX;
	Rect crtc_rects[2] = {};
	bool skip_shift_x = false, skip_shift_y = false;

	if (EN1 && EN2 && anti_blur)
	{
		if (pr.FBP(0) == pr.FBP(1) && pr.PSM(0) == pr.PSM(1) && pr.FBW(0) == pr.FBW(1) && pr.DBX(0) == pr.DBX(1) &&
		    pr.MAGH(0) == pr.MAGH(1) && pr.MAGV(0) == pr.MAGV(1))
		{
			if (MMOD == 1 && pr.DY(0) == pr.DY(1) && pr.DBY(0) == pr.DBY(1) && pr.DBX(0) == pr.DBX(1) &&
			    pr.DX(0) != pr.DX(1) && std::abs(int(pr.DX(0)) - int(pr.DX(1))) < 2 * int(pr.MAGH(0) + 1))
			{
				EN2 = false;
				ALP = 0xff;
				MMOD = 1;
			}
			else if (alternative_sampling && pr.DY(0) == pr.DY(1))
			{
				if (pr.DBY(0) + 1 == pr.DBY(1))
				{
					EN2 = false;
					ALP = 0xff;
					MMOD = 1;
				}
				else if (pr.DBY(0) == pr.DBY(1) + 1)
				{
					EN1 = false;
					SLBG = 0;
				}
			}
			else if (alternative_sampling && pr.DBY(0) == pr.DBY(1))
			{
				if (pr.DY(0) + 1 == pr.DY(1))
				{
					EN2 = false;
					ALP = 0xff;
					MMOD = 1;
				}
				else if (pr.DY(0) == pr.DY(1) + 1)
				{
					EN1 = false;
					SLBG = 0;
				}
			}
		}
	}

	std::shared_ptr<ScanImage> circuit[2];
	const bool enabled[2] = { EN1, EN2 };
	for (int c = 0; c < 2; c++)
	{
		if (!enabled[c])
			continue;
		SamplingRect rect = compute_circuit_rect(pr, phase, c, force_progressive);
		uint32_t w = rect.image_w, h = rect.image_h;
		if (w && h)
		{
			if (high_resolution_scanout)
			{
				w *= 2;
				h *= 2;
			}
			circuit[c] = sample_circuit(vram.data(), pr, c, rect, w, h, super_samples);
		}
		int off_x = int(pr.DX(c) + clock_divider / 2) / int(clock_divider) - scan_offset_x;
		int off_y = int(pr.DY(c) + rect.phase_offset) >> int(is_interlaced);
		off_y -= scan_offset_y;
		uint32_t width = (pr.DW(c) + 1) / clock_divider;
		uint32_t height = (pr.DH(c) + 1 + uint32_t(is_interlaced)) >> uint32_t(is_interlaced);
		if (!is_interlaced)
			height = (height + 1) & ~1u;
		crtc_rects[c] = { off_x, off_y, width, height };
		if (width != mode_width)
			skip_shift_x = true;
		if (height != mode_height)
			skip_shift_y = true;
		crtc_shift_x = std::min(off_x, crtc_shift_x);
		crtc_shift_y = std::min(off_y, crtc_shift_y);
	}

	if (!info.overscan && !info.crtc_offsets)
	{
		for (int c = 0; c < 2; c++)
		{
			if (!enabled[c])
				continue; // the rectangle of a disabled circuit is never drawn
			if (!skip_shift_x)
				crtc_rects[c].x -= crtc_shift_x;
			if (!skip_shift_y)
				crtc_rects[c].y -= crtc_shift_y;
		}
	}

	uint32_t real_mode_width = mode_width;
	float sampling_rate_mhz = 54.0f / float(clock_divider);
	if (clock_divider == 4 && !is_interlaced)
		sampling_rate_mhz *= 2.0f;

	if (info.adapt_to_internal_horizontal_resolution)
	{
		uint32_t hres0 = circuit[0] ? circuit[0]->w : 0;
		uint32_t hres1 = circuit[1] ? circuit[1]->w : 0;
		if (high_resolution_scanout)
		{
			hres0 >>= 1;
			hres1 >>= 1;
		}
		uint32_t magh1 = pr.MAGH(0) + 1, magh2 = pr.MAGH(1) + 1;
		if (hres0 == 0)
		{
			hres0 = hres1;
			magh1 = magh2;
		}
		if (hres1 == 0)
			hres1 = hres0;
		if (hres0 && hres0 == hres1)
		{
			float width_scaling = float(clock_divider) / float(magh1);
			sampling_rate_mhz *= width_scaling;
			for (int c = 0; c < 2; c++)
			{
				crtc_rects[c].x = int32_t(std::round(float(crtc_rects[c].x) * width_scaling));
				crtc_rects[c].w = uint32_t(std::round(float(crtc_rects[c].w) * width_scaling));
			}
			mode_width = uint32_t(std::round(float(mode_width) * width_scaling));
		}
	}

	out.mode_width = mode_width;
	out.mode_height = mode_height;
	out.high_resolution_scanout = high_resolution_scanout;
	out.sampling_rate_mhz = sampling_rate_mhz;

	if (info.raw_circuit_scanout && !info.crtc_offsets && !info.overscan &&
	    info.adapt_to_internal_horizontal_resolution && !force_deinterlace && !is_interlaced && !pr.WRITE())
	{
		uint32_t emw = mode_width, emh = mode_height;
		if (high_resolution_scanout)
		{
			emw *= 2;
			emh *= 2;
		}
		bool raw1 = circuit[0] && !circuit[1] && MMOD == 1 && ALP == 0xff && circuit[0]->w <= emw &&
		            circuit[0]->h <= emh;
		bool raw2 = circuit[1] && !circuit[0] && SLBG == 0 && circuit[1]->w <= emw && circuit[1]->h <= emh;
		std::shared_ptr<ScanImage> img = raw1 ? circuit[0] : raw2 ? circuit[1] : nullptr;
		if (img)
		{
			emit(img, out, image);
			out.internal_width = img->w >> int(high_resolution_scanout);
			out.internal_height = img->h >> int(high_resolution_scanout);
			out.double_strike = double_strike;
			return 0;
		}
	}

	out.internal_width = mode_width;
	out.internal_height = mode_height;
	uint32_t mw = mode_width << int(high_resolution_scanout), mh = mode_height << int(high_resolution_scanout);
	if (field_aware_rendering)
		mh--;
	auto merged = std::make_shared<ScanImage>(mw, mh);
	for (auto &p : merged->px)
	{
		float bg[4] = { float(pr.BG(0)) * (1.0f / 255.0f), float(pr.BG(1)) * (1.0f / 255.0f),
		                float(pr.BG(2)) * (1.0f / 255.0f), 0.0f };
		p = pack_unorm(bg);
	}
	auto viewport = [&](const Rect &r, float v[4]) {
		v[0] = float(r.x);
		v[1] = float(r.y);
		v[2] = float(r.w);
		v[3] = float(r.h);
		if (high_resolution_scanout)
		{
			for (int i = 0; i < 4; i++)
				v[i] *= 2.0f;
			if (field_aware_rendering && !info.phase)
				v[1] -= 1.0f;
		}
	};
	if (SLBG == 0 && circuit[1] && crtc_rects[1].w && crtc_rects[1].h)
	{
		float v[4];
		viewport(crtc_rects[1], v);
		blit(*merged, *circuit[1], v[0], v[1], v[2], v[3], 0, 0.0f);
	}
	if (circuit[0] && crtc_rects[0].w && crtc_rects[0].h)
	{
		float v[4];
		viewport(crtc_rects[0], v);
		int mode = MMOD == 1 ? (ALP != 0xff ? 1 : 0) : 2;
		float alp = float(pr.ALP()) * (1.0f / 255.0f);
		blit(*merged, *circuit[0], v[0], v[1], v[2], v[3], mode, alp);
	}

	if (pr.WRITE())
	{
		uint32_t res_x = pr.WW() + 1, res_y = pr.WH() + 1;
		if (!pr.WFFMD())
			res_y *= 2;
		const ScanImage *view = pr.FBIN() == 0 ? merged.get() : circuit[1] ? circuit[1].get() : circuit[0].get();
		if (view)
		{
			uint32_t scanout_width = real_mode_width, scanout_height = merged->h;
			uint32_t max_w = (scanout_width * clock_divider + pr.SMPH()) / (pr.SMPH() + 1);
			uint32_t max_h = (scanout_height + pr.SMPV()) / (pr.SMPV() + 1);
			res_x = std::min(res_x, max_w);
			res_y = std::min(res_y, max_h);
			float uv_base[2] = { 0.5f / float(res_x), 0.5f / float(res_y) };
			float uv_scale[2] = { float(pr.SMPH() + 1) / float(scanout_width * clock_divider),
			                      float(pr.SMPV() + 1) / float(scanout_height) };
			extwrite(vram.data(), sample_log2 != 0, *view, res_x, res_y, pr.WDX(), pr.WDY(), uv_base, uv_scale,
			         pr.EXBP(), pr.EXBW(), pr.EMODA(), pr.EMODC());
		}
	}

	bool should_deinterlace = !high_resolution_scanout && (is_interlaced || force_deinterlace);
	if (!info.skip_deinterlace && should_deinterlace)
	{
		if (last_fields[0] && (last_fields[0]->w != merged->w || last_fields[0]->h != merged->h))
			for (auto &f : last_fields)
				f.reset();
		for (int i = 3; i >= 1; i--)
			last_fields[i] = std::move(last_fields[i - 1]);
		last_fields[0] = merged;
		if (!last_fields[1])
			last_fields[1] = last_fields[0];
		if (!last_fields[2])
			last_fields[2] = last_fields[0];
		if (!last_fields[3])
			last_fields[3] = last_fields[1];
		merged = weave(last_fields, info.phase);
	}
	else
	{
		for (auto &f : last_fields)
			f.reset();
	}

	emit(merged, out, image);
	out.interlaced = should_deinterlace;
	out.interlace_phase = info.phase;
	out.double_strike = double_strike;
	return 0;
}
}

extern "C"
{
void orc_set_priv(void *g, const void *data, size_t size)
{
	auto *gs = static_cast<orc::GS *>(g);
	memcpy(gs->priv.data(), data, std::min(size, gs->priv.size()));
}
uint32_t orc_complex_scanmsk_timeout(void *g) { return static_cast<orc::GS *>(g)->complex_scanmsk_timeout; }
int orc_vsync_can_skip(void *g, const orc::VsyncInfo *info) { return static_cast<orc::GS *>(g)->vsync_can_skip(*info); }
// Scans out; writes up to `cap` pixels of the image to `pixels`. Returns the vsync status.
int orc_vsync(void *g, const orc::VsyncInfo *info, orc::ScanoutOut *out, uint32_t *pixels, size_t cap)
{
	std::vector<uint32_t> image;
	int rc = static_cast<orc::GS *>(g)->vsync(*info, *out, image);
	if (pixels)
		memcpy(pixels, image.data(), std::min(cap, image.size()) * 4);
	return rc;
}
}
