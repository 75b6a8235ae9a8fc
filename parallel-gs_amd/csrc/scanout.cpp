// CRTC scanout: host decisions of the reference's GSRenderer::vsync and the launches of the
// scanout kernels. See scanout.hpp.
#include "scanout.hpp"
#include "pgs_kernels.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace pgs
{
namespace
{
uint64_t reg64(const uint8_t *priv, uint32_t offset)
{
	uint64_t v;
	memcpy(&v, priv + offset, sizeof(v));
	return v;
}

uint32_t bits(uint64_t v, uint32_t lo, uint32_t n)
{
	return uint32_t((v >> lo) & ((uint64_t(1) << n) - 1u));
}

// SMODE1.CMOD values and LC line counts (gs/gs_registers.hpp:711-726).
constexpr uint32_t kCmodProgressive = 0, kCmodNtsc = 2, kCmodPal = 3;
constexpr uint32_t kLcAnalog = 32, kLcHdtv = 22, kLcVga = 15, kLcUnknown29 = 29;
constexpr uint32_t kDividerComposite = 4, kDividerComponent = 2, kDividerHdtv = 1;

// compute_circuit_rect (gs/gs_renderer.cpp:4114-4172): the sampled extent of one read circuit,
// and where interlaced output starts and steps through its lines.
struct SamplingRect
{
	uint32_t image_w, image_h, valid_w, valid_h, phase_offset, phase_stride;
};

SamplingRect circuit_rect(const CrtcRegisters &r, uint32_t phase, uint32_t c, bool force_progressive)
{
	bool is_interlaced = r.interlace != 0;
	const bool alternative = is_interlaced && !r.ffmd;
	if (alternative && force_progressive)
		is_interlaced = false;
	const auto &d = r.display[c];
	SamplingRect s{};
	s.image_w = (d.dw + 1u) / (d.magh + 1u);
	s.image_h = (d.dh + 1u) / (d.magv + 1u);
	s.valid_w = s.image_w;
	s.phase_offset = 0;
	s.phase_stride = 1;
	if (is_interlaced)
	{
		if (alternative)
		{
			s.phase_offset = (d.dy ^ phase) & 1u;
			s.phase_stride = 2;
		}
		s.valid_h = (s.image_h - s.phase_offset + 1u) >> 1;
		s.image_h = (s.image_h + 1u) >> 1;
	}
	else
	{
		s.valid_h = s.image_h;
		s.image_h = (s.image_h + 1u) & ~1u;
	}
	return s;
}

int32_t round_scaled(int32_t v, float scale)
{
	return int32_t(std::round(float(v) * scale));
}
} // namespace

CrtcRegisters decode_crtc_registers(const uint8_t *priv)
{
	CrtcRegisters r{};
	const uint64_t pmode = reg64(priv, 0x00), smode1 = reg64(priv, 0x10), smode2 = reg64(priv, 0x20);
	r.en[0] = bits(pmode, 0, 1);
	r.en[1] = bits(pmode, 1, 1);
	r.mmod = bits(pmode, 5, 1);
	r.slbg = bits(pmode, 7, 1);
	r.alp = bits(pmode, 8, 8);
	r.lc = bits(smode1, 3, 7);
	r.cmod = bits(smode1, 13, 2);
	r.interlace = bits(smode2, 0, 1);
	r.ffmd = bits(smode2, 1, 1);
	for (uint32_t c = 0; c < 2; c++)
	{
		const uint64_t fb = reg64(priv, 0x70 + 0x20 * c), dp = reg64(priv, 0x80 + 0x20 * c);
		r.dispfb[c] = { bits(fb, 0, 9), bits(fb, 9, 6), bits(fb, 15, 5), bits(fb, 32, 11), bits(fb, 43, 11) };
		r.display[c] = { bits(dp, 0, 12), bits(dp, 12, 11), bits(dp, 23, 4), bits(dp, 27, 2), bits(dp, 32, 12),
		                 bits(dp, 44, 11) };
	}
	const uint64_t extbuf = reg64(priv, 0xb0), extdata = reg64(priv, 0xc0);
	r.exbp = bits(extbuf, 0, 14);
	r.exbw = bits(extbuf, 14, 6);
	r.fbin = bits(extbuf, 20, 2);
	r.wffmd = bits(extbuf, 22, 1);
	r.emoda = bits(extbuf, 23, 2);
	r.emodc = bits(extbuf, 25, 2);
	r.wdx = bits(extbuf, 32, 11);
	r.wdy = bits(extbuf, 43, 11);
	r.smph = bits(extdata, 23, 4);
	r.smpv = bits(extdata, 27, 2);
	r.ww = bits(extdata, 32, 12);
	r.wh = bits(extdata, 44, 11);
	r.extwrite = bits(reg64(priv, 0xd0), 0, 1);
	r.bgcolor = uint32_t(reg64(priv, 0xe0) & 0xffffffu);
	return r;
}

bool scanout_is_interlaced(const CrtcRegisters &r, const VSyncInfo &info)
{
	bool is_interlaced = r.interlace != 0;
	if (is_interlaced && !r.ffmd && info.force_progressive)
		is_interlaced = false;
	return is_interlaced;
}

// GSRenderer::vsync (gs/gs_renderer.cpp:4189-5055), the decisions only.
ScanoutPlan plan_scanout(const CrtcRegisters &r, const VSyncInfo &info, uint32_t rate_x_log2, uint32_t rate_y_log2)
{
	ScanoutPlan p{};
	const bool progressive_cmod = r.cmod == kCmodProgressive;
	bool hr = info.high_resolution_scanout;
	p.interlaced = scanout_is_interlaced(r, info);
	p.force_deinterlace = !hr && r.ffmd && r.interlace && !progressive_cmod;
	const bool alternative = r.interlace && !r.ffmd;
	p.double_strike = !r.interlace && !progressive_cmod;
	if (!info.force_progressive || r.extwrite || !rate_x_log2 || !rate_y_log2 || p.force_deinterlace)
		hr = false;
	uint32_t super_samples = hr ? 1u << (rate_x_log2 + rate_y_log2) : 1u;

	// Video mode table (gs/gs_renderer.cpp:4253-4368).
	uint32_t divider = 1, mode_w = 0, mode_h = 0;
	int32_t scan_x = 0, scan_y = 0;
	const bool tall = !p.interlaced && !p.double_strike && !p.force_deinterlace;
	if ((r.cmod == kCmodNtsc || (progressive_cmod && !r.interlace)) && r.lc == kLcAnalog)
	{
		if (info.overscan) { mode_w = 712; mode_h = 240; scan_x = 123; scan_y = 17; }
		else { mode_w = 640; mode_h = 224; scan_x = 159; scan_y = 25; }
		if (tall) { mode_h *= 2; scan_y *= 2; }
		divider = progressive_cmod ? kDividerComponent : kDividerComposite;
	}
	else if (r.cmod == kCmodPal && r.lc == kLcAnalog)
	{
		if (info.overscan) { mode_w = 712; mode_h = 288; scan_x = 133; scan_y = 21; }
		else { mode_w = 640; mode_h = 256; scan_x = 169; scan_y = 36; }
		if (tall) { mode_h *= 2; scan_y *= 2; }
		divider = kDividerComposite;
	}
	else if (progressive_cmod && r.lc == kLcHdtv)
	{
		if (r.interlace) { mode_w = 1920; mode_h = 540; scan_x = 236; scan_y = 0; }
		else { mode_w = 1280; mode_h = 720; scan_x = 300; scan_y = 0; }
		if (tall) { mode_h *= 2; scan_y *= 2; }
		divider = kDividerHdtv;
		hr = false;
		super_samples = 1;
	}
	else if (progressive_cmod && r.lc == kLcVga)
	{
		mode_w = 640; mode_h = 480; scan_x = 138; scan_y = 34;
		divider = kDividerComponent;
	}
	else if (r.lc == kLcUnknown29)
	{
		mode_w = 640; mode_h = 480; scan_x = 318; scan_y = 50;
		divider = kDividerComposite;
	}
	else
		return p; // unknown video format: no frame

	p.known_mode = true;
	p.high_resolution = hr;
	p.super_samples = super_samples;
	p.clock_divider = divider;
	const uint32_t phase = (alternative && !info.force_progressive) ? info.phase : 0u;

	// Anti-blur (gs/gs_renderer.cpp:4478-4545): two circuits showing the same buffer one line
	// or a fraction of a pixel apart are reduced to one.
	bool en[2] = { r.en[0] != 0, r.en[1] != 0 };
	uint32_t mmod = r.mmod, alp = r.alp, slbg = r.slbg;
	if (en[0] && en[1] && info.anti_blur)
	{
		const auto &f1 = r.dispfb[0], &f2 = r.dispfb[1];
		const auto &d1 = r.display[0], &d2 = r.display[1];
		auto keep_circuit1 = [&] { en[1] = false; alp = 0xff; mmod = 1; };
		auto keep_circuit2 = [&] { en[0] = false; slbg = 0; };
		if (f1.fbp == f2.fbp && f1.psm == f2.psm && f1.fbw == f2.fbw && f1.dbx == f2.dbx && d1.magh == d2.magh &&
		    d1.magv == d2.magv)
		{
			if (mmod == 1 && d1.dy == d2.dy && f1.dby == f2.dby && f1.dbx == f2.dbx && d1.dx != d2.dx &&
			    std::abs(int32_t(d1.dx) - int32_t(d2.dx)) < int32_t(2 * (d1.magh + 1)))
				keep_circuit1();
			else if (alternative && d1.dy == d2.dy)
			{
				if (f1.dby + 1 == f2.dby)
					keep_circuit1();
				else if (f1.dby == f2.dby + 1)
					keep_circuit2();
			}
			else if (alternative && f1.dby == f2.dby)
			{
				if (d1.dy + 1 == d2.dy)
					keep_circuit1();
				else if (d1.dy == d2.dy + 1)
					keep_circuit2();
			}
		}
	}
	p.mmod = mmod;
	p.alp = alp;
	p.slbg = slbg;
	p.bgcolor = r.bgcolor;

	// Read circuits and their CRTC rectangles (gs/gs_renderer.cpp:4547-4657).
	int32_t shift_x = INT32_MAX, shift_y = INT32_MAX;
	bool skip_shift_x = false, skip_shift_y = false;
	for (uint32_t c = 0; c < 2; c++)
	{
		if (!en[c])
			continue;
		auto &pc = p.circuit[c];
		const SamplingRect s = circuit_rect(r, phase, c, info.force_progressive);
		if (s.image_w && s.image_h)
		{
			pc.image = true;
			const uint32_t k = hr ? 2u : 1u;
			pc.width = s.image_w * k;
			pc.height = s.image_h * k;
			const uint32_t kv = super_samples > 1u ? 2u : 1u;
			pc.valid_width = s.valid_w * kv;
			pc.valid_height = s.valid_h * kv;
			pc.psm = r.dispfb[c].psm;
			pc.fbp = r.dispfb[c].fbp;
			pc.fbw = r.dispfb[c].fbw;
			pc.dbx = r.dispfb[c].dbx;
			pc.dby = r.dispfb[c].dby;
			pc.phase_offset = s.phase_offset;
			pc.phase_stride = s.phase_stride;
		}
		const auto &d = r.display[c];
		const int32_t off_x = int32_t(d.dx + divider / 2u) / int32_t(divider) - scan_x;
		const int32_t off_y = (int32_t(d.dy + s.phase_offset) >> int32_t(p.interlaced)) - scan_y;
		const uint32_t width = (d.dw + 1u) / divider;
		uint32_t height = (d.dh + 1u + uint32_t(p.interlaced)) >> uint32_t(p.interlaced);
		if (!p.interlaced)
			height = (height + 1u) & ~1u;
		pc.x = off_x;
		pc.y = off_y;
		pc.w = width;
		pc.h = height;
		skip_shift_x |= width != mode_w;
		skip_shift_y |= height != mode_h;
		shift_x = std::min(shift_x, off_x);
		shift_y = std::min(shift_y, off_y);
	}
	if (!info.overscan && !info.crtc_offsets)
	{
		for (uint32_t c = 0; c < 2; c++)
		{
			if (!en[c])
				continue;
			if (!skip_shift_x)
				p.circuit[c].x -= shift_x;
			if (!skip_shift_y)
				p.circuit[c].y -= shift_y;
		}
	}

	p.real_mode_width = mode_w;
	float rate_mhz = 54.0f / float(divider);
	if (divider == kDividerComposite && !p.interlaced)
		rate_mhz *= 2.0f;

	// Horizontal resolution adapted to MAGH (gs/gs_renderer.cpp:4675-4709).
	if (info.adapt_to_internal_horizontal_resolution)
	{
		uint32_t h0 = p.circuit[0].image ? p.circuit[0].width : 0u;
		uint32_t h1 = p.circuit[1].image ? p.circuit[1].width : 0u;
		if (hr)
		{
			h0 >>= 1;
			h1 >>= 1;
		}
		uint32_t magh1 = r.display[0].magh + 1u, magh2 = r.display[1].magh + 1u;
		if (h0 == 0)
		{
			h0 = h1;
			magh1 = magh2;
		}
		if (h1 == 0)
			h1 = h0;
		if (h0 && h0 == h1)
		{
			const float scale = float(divider) / float(magh1);
			rate_mhz *= scale;
			for (uint32_t c = 0; c < 2; c++)
			{
				p.circuit[c].x = round_scaled(p.circuit[c].x, scale);
				p.circuit[c].w = uint32_t(std::round(float(p.circuit[c].w) * scale));
			}
			mode_w = uint32_t(std::round(float(mode_w) * scale));
		}
	}
	p.mode_width = mode_w;
	p.mode_height = mode_h;
	p.sampling_rate_mhz = rate_mhz;

	// Raw circuit shortcut (gs/gs_renderer.cpp:4717-4757).
	if (info.raw_circuit_scanout && !info.crtc_offsets && !info.overscan &&
	    info.adapt_to_internal_horizontal_resolution && !p.force_deinterlace && !p.interlaced && !r.extwrite)
	{
		const uint32_t ew = mode_w << uint32_t(hr), eh = mode_h << uint32_t(hr);
		const auto &c1 = p.circuit[0], &c2 = p.circuit[1];
		if (c1.image && !c2.image && mmod == 1 && alp == 0xff && c1.width <= ew && c1.height <= eh)
			p.raw_circuit = 0;
		else if (c2.image && !c1.image && slbg == 0 && c2.width <= ew && c2.height <= eh)
			p.raw_circuit = 1;
		if (p.raw_circuit >= 0)
			return p;
	}

	p.merged_width = mode_w << uint32_t(hr);
	p.merged_height = mode_h << uint32_t(hr);

	// EXTWRITE feedback (gs/gs_renderer.cpp:4921-4999).
	if (r.extwrite)
	{
		if (r.fbin == 0)
			p.extwrite_view = -1;
		else if (p.circuit[1].image)
			p.extwrite_view = 1;
		else if (p.circuit[0].image)
			p.extwrite_view = 0;
		else
			p.extwrite_view = -2; // no source circuit: nothing is written
		if (p.extwrite_view != -2)
		{
			const uint32_t scan_w = p.real_mode_width, scan_h = p.merged_height;
			uint32_t w = r.ww + 1u, h = r.wh + 1u;
			if (!r.wffmd)
				h *= 2u;
			const uint32_t max_w = (scan_w * divider + r.smph) / (r.smph + 1u);
			const uint32_t max_h = (scan_h + r.smpv) / (r.smpv + 1u);
			p.ext_width = std::min(w, max_w);
			p.ext_height = std::min(h, max_h);
			p.extwrite = p.ext_width && p.ext_height;
			if (p.extwrite)
			{
				p.wdx = r.wdx;
				p.wdy = r.wdy;
				p.exbp = r.exbp;
				p.exbw = r.exbw;
				p.emoda = r.emoda;
				p.emodc = r.emodc;
				p.uv_base[0] = 0.5f / float(p.ext_width);
				p.uv_base[1] = 0.5f / float(p.ext_height);
				p.uv_scale[0] = float(r.smph + 1u) / float(scan_w * divider);
				p.uv_scale[1] = float(r.smpv + 1u) / float(scan_h);
			}
		}
	}
	p.should_deinterlace = !hr && (p.interlaced || p.force_deinterlace);
	return p;
}

// ---------------------------------------------------------------------------------------

Scanout::~Scanout()
{
	for (auto &i : circuit_)
		(void)hipFree(i.ptr);
	for (auto &i : pool_)
		(void)hipFree(i.ptr);
	(void)hipFree(deinterlaced_.ptr);
	for (auto &t : timed_)
	{
		(void)hipEventDestroy(t.start);
		(void)hipEventDestroy(t.stop);
	}
}

int Scanout::ensure(Image &img, size_t pixels)
{
	if (img.cap >= pixels)
		return 0;
	(void)hipFree(img.ptr);
	img.ptr = nullptr;
	img.cap = 0;
	hipError_t e = hipMalloc(reinterpret_cast<void **>(&img.ptr), pixels * 4u);
	if (e != hipSuccess)
		return -int(e);
	img.cap = pixels;
	return 0;
}

void Scanout::reset_history()
{
	for (int &h : history_)
		h = -1;
	history_w_ = history_h_ = 0;
}

double Scanout::poll_ms(bool wait)
{
	double ms = 0.0;
	size_t done = 0;
	for (; done < timed_.size(); done++)
	{
		Timed &t = timed_[done];
		if (wait)
			(void)hipEventSynchronize(t.stop);
		else if (hipEventQuery(t.stop) != hipSuccess)
			break;
		float e = 0.0f;
		if (hipEventElapsedTime(&e, t.start, t.stop) == hipSuccess)
			ms += e;
		(void)hipEventDestroy(t.start);
		(void)hipEventDestroy(t.stop);
	}
	timed_.erase(timed_.begin(), timed_.begin() + std::ptrdiff_t(done));
	return ms;
}

int Scanout::run(const ScanoutPlan &p, const VSyncInfo &info, uint8_t *vram, uint32_t sample_log2,
                 hipStream_t stream, ScanoutResult &res)
{
	res = ScanoutResult{};
	if (!p.known_mode)
		return 0;
	Timed t{ nullptr, nullptr };
	if (timing_)
	{
		if (hipEventCreate(&t.start) != hipSuccess || hipEventCreate(&t.stop) != hipSuccess)
			return -int(hipErrorOutOfMemory);
		(void)hipEventRecord(t.start, stream);
	}
	int rc = 0;
	auto ok = [&](int v) { if (rc == 0 && v != 0) rc = v; return rc == 0; };

	for (uint32_t c = 0; c < 2; c++)
	{
		const auto &pc = p.circuit[c];
		if (!pc.image)
			continue;
		if (!ok(ensure(circuit_[c], size_t(pc.width) * pc.height)))
			break;
		pgs_scanout_circuit_args a{};
		a.vram = vram;
		a.image = circuit_[c].ptr;
		a.width = pc.width;
		a.height = pc.height;
		a.valid_width = pc.valid_width;
		a.valid_height = pc.valid_height;
		a.psm = pc.psm;
		a.base_block = pc.fbp * 32u;
		a.fbw = pc.fbw;
		a.dbx = pc.dbx;
		a.dby = pc.dby;
		a.phase = pc.phase_offset;
		a.phase_stride = pc.phase_stride;
		a.super_samples = p.super_samples;
		ok(pgs_k_scanout_circuit(&a, stream));
	}

	res.mode_width = p.mode_width;
	res.mode_height = p.mode_height;
	res.high_resolution_scanout = p.high_resolution;
	res.sampling_rate_mhz = p.sampling_rate_mhz;
	res.double_strike = p.double_strike;

	if (rc == 0 && p.raw_circuit >= 0)
	{
		const auto &pc = p.circuit[p.raw_circuit];
		res.image = circuit_[p.raw_circuit].ptr;
		res.width = pc.width;
		res.height = pc.height;
		res.internal_width = pc.width >> uint32_t(p.high_resolution);
		res.internal_height = pc.height >> uint32_t(p.high_resolution);
	}
	else if (rc == 0)
	{
		res.internal_width = p.mode_width;
		res.internal_height = p.mode_height;
		const uint32_t mw = p.merged_width, mh = p.merged_height;
		const bool deinterlace = !info.skip_deinterlace && p.should_deinterlace;
		// A field of another size starts a new history (the reference would sample fields of
		// different extents together).
		if (!deinterlace || mw != history_w_ || mh != history_h_)
			reset_history();
		int slot = 0;
		for (; slot < 4; slot++)
			if (slot != history_[0] && slot != history_[1] && slot != history_[2])
				break;
		if (ok(ensure(pool_[slot], size_t(mw) * mh)))
		{
			uint32_t *merged = pool_[slot].ptr;
			pgs_scanout_merge_args m{};
			m.merged = merged;
			m.width = mw;
			m.height = mh;
			const uint32_t k = p.high_resolution ? 2u : 1u;
			for (uint32_t c = 0; c < 2; c++)
			{
				const auto &pc = p.circuit[c];
				m.circuit[c] = circuit_[c].ptr;
				m.circuit_width[c] = pc.width;
				m.circuit_height[c] = pc.height;
				m.vp_x[c] = pc.x * int32_t(k);
				m.vp_y[c] = pc.y * int32_t(k);
				m.vp_width[c] = pc.w * k;
				m.vp_height[c] = pc.h * k;
			}
			if (p.slbg == 0 && p.circuit[1].image && p.circuit[1].w && p.circuit[1].h)
				m.draw |= 2u;
			if (p.circuit[0].image && p.circuit[0].w && p.circuit[0].h)
				m.draw |= 1u;
			m.blend = p.mmod == 1 ? (p.alp != 0xffu ? 1u : 0u) : 2u;
			m.alp = p.alp;
			m.bgcolor = p.bgcolor;
			ok(pgs_k_scanout_merge(&m, stream));
			res.image = merged;
			res.width = mw;
			res.height = mh;

			if (rc == 0 && p.extwrite)
			{
				pgs_scanout_extwrite_args e{};
				e.vram = vram;
				if (p.extwrite_view < 0)
				{
					e.view = merged;
					e.view_width = mw;
					e.view_height = mh;
				}
				else
				{
					e.view = circuit_[p.extwrite_view].ptr;
					e.view_width = p.circuit[p.extwrite_view].width;
					e.view_height = p.circuit[p.extwrite_view].height;
				}
				e.width = p.ext_width;
				e.height = p.ext_height;
				e.wdx = p.wdx;
				e.wdy = p.wdy;
				e.exbp = p.exbp;
				e.exbw = p.exbw;
				e.emoda = p.emoda;
				e.emodc = p.emodc;
				for (int i = 0; i < 2; i++)
				{
					e.uv_base[i] = p.uv_base[i];
					e.uv_scale[i] = p.uv_scale[i];
				}
				e.clear_reference = sample_log2 != 0;
				ok(pgs_k_scanout_extwrite(&e, stream));
			}

			if (rc == 0 && deinterlace)
			{
				for (int i = 3; i >= 1; i--)
					history_[i] = history_[i - 1];
				history_[0] = slot;
				if (history_[1] < 0)
					history_[1] = history_[0];
				if (history_[2] < 0)
					history_[2] = history_[0];
				if (history_[3] < 0)
					history_[3] = history_[1];
				history_w_ = mw;
				history_h_ = mh;
				if (ok(ensure(deinterlaced_, size_t(mw) * mh * 2u)))
				{
					pgs_scanout_weave_args w{};
					for (int i = 0; i < 4; i++)
						w.field[i] = pool_[history_[i]].ptr;
					w.out = deinterlaced_.ptr;
					w.width = mw;
					w.height = mh * 2u;
					w.phase = info.phase;
					ok(pgs_k_scanout_weave(&w, stream));
					res.image = deinterlaced_.ptr;
					res.height = mh * 2u;
				}
			}
			res.interlaced = p.should_deinterlace;
			res.interlace_phase = info.phase;
		}
	}
	if (timing_)
	{
		(void)hipEventRecord(t.stop, stream);
		timed_.push_back(t);
	}
	if (rc != 0)
		res = ScanoutResult{};
	return rc;
}
}
