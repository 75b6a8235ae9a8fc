// CRTC scanout (vsync): turns the privileged display registers and VRAM into a frame.
//
// Host half of the reference's GSRenderer::vsync (gs/gs_renderer.cpp:4114-5055): the video
// mode table, the read-circuit rectangles, CRTC offsets, anti-blur, the raw-circuit shortcut,
// the EXTWRITE set-up and the deinterlacer's field history are decided here; the pixels are
// produced by the kernels of kernels_scanout.hip on the renderer's stream.
#pragma once
#include "gs_defs.hpp"
#include <cstdint>
#include <vector>

namespace pgs
{
// VSyncInfo (gs/gs_interface.hpp:155-198) without the Vulkan layout / stage / access fields:
// the result is a device buffer ordered on the context's stream.
struct VSyncInfo
{
	uint32_t phase = 0;
	bool force_progressive = false;
	bool overscan = false;
	bool anti_blur = false;
	bool crtc_offsets = false;
	bool adapt_to_internal_horizontal_resolution = false;
	bool raw_circuit_scanout = false;
	bool high_resolution_scanout = false;
	bool skip_deinterlace = false;
};

// ScanoutResult (gs/gs_renderer.hpp:23-52). image is NULL when the video mode is unknown.
struct ScanoutResult
{
	const uint32_t *image = nullptr; // RGBA8, row pitch = width
	uint32_t width = 0, height = 0;
	uint32_t internal_width = 0, internal_height = 0;
	uint32_t mode_width = 0, mode_height = 0;
	bool high_resolution_scanout = false;
	bool interlaced = false;
	bool double_strike = false;
	uint32_t interlace_phase = 0;
	float sampling_rate_mhz = 0.0f;
};

// The privileged registers the scanout reads, decoded from the 8 KiB register block (one
// 16-byte slot per register: PMODE 0x00, SMODE1 0x10, SMODE2 0x20, DISPFB1 0x70, DISPLAY1 0x80,
// DISPFB2 0x90, DISPLAY2 0xa0, EXTBUF 0xb0, EXTDATA 0xc0, EXTWRITE 0xd0, BGCOLOR 0xe0;
// gs/gs_interface.hpp:70-108, bit layouts gs/gs_registers.hpp:562-736).
struct CrtcRegisters
{
	uint32_t en[2], mmod, slbg, alp;
	uint32_t cmod, lc;
	uint32_t interlace, ffmd;
	struct { uint32_t fbp, fbw, psm, dbx, dby; } dispfb[2];
	struct { uint32_t dx, dy, magh, magv, dw, dh; } display[2];
	uint32_t exbp, exbw, fbin, wffmd, emoda, emodc, wdx, wdy;
	uint32_t smph, smpv, ww, wh;
	uint32_t extwrite;
	uint32_t bgcolor; // R | G << 8 | B << 16
};
CrtcRegisters decode_crtc_registers(const uint8_t *priv);

// scanout_is_interlaced (gs/gs_renderer.cpp:4174-4182).
bool scanout_is_interlaced(const CrtcRegisters &r, const VSyncInfo &info);

// Everything the kernels need for one vsync, computed on the host.
struct ScanoutPlan
{
	bool known_mode = false;
	bool high_resolution = false;
	bool interlaced = false, force_deinterlace = false, double_strike = false;
	uint32_t super_samples = 1;
	uint32_t clock_divider = 1;
	uint32_t mode_width = 0, mode_height = 0, real_mode_width = 0;
	float sampling_rate_mhz = 0.0f;
	uint32_t mmod = 0, alp = 0, slbg = 0;
	uint32_t bgcolor = 0;
	struct Circuit
	{
		bool image = false;  // an image is sampled (enabled, non-empty extent)
		uint32_t psm = 0, fbp = 0, fbw = 0, dbx = 0, dby = 0, phase_offset = 0, phase_stride = 1;
		uint32_t width = 0, height = 0, valid_width = 0, valid_height = 0;
		int32_t x = 0, y = 0;  // CRTC rectangle in the merged frame (single-sampled units)
		uint32_t w = 0, h = 0;
	} circuit[2];
	int raw_circuit = -1;     // circuit returned as the frame (raw_circuit_scanout), or -1
	uint32_t merged_width = 0, merged_height = 0;
	bool extwrite = false;
	int extwrite_view = -1;    // -1 merged frame, else circuit index
	uint32_t ext_width = 0, ext_height = 0;
	uint32_t wdx = 0, wdy = 0, exbp = 0, exbw = 0, emoda = 0, emodc = 0;
	float uv_base[2] = {}, uv_scale[2] = {};
	bool should_deinterlace = false;
};

ScanoutPlan plan_scanout(const CrtcRegisters &r, const VSyncInfo &info, uint32_t sampling_rate_x_log2,
                         uint32_t sampling_rate_y_log2);

class Renderer;

// Device images of the scanout and the deinterlacer's history of merged fields.
class Scanout
{
public:
	Scanout() = default;
	~Scanout();
	Scanout(const Scanout &) = delete;
	Scanout &operator=(const Scanout &) = delete;

	// Runs the kernels of a plan on `stream` over `vram` (all planes). Returns 0 or a negative
	// hipError_t. The images stay valid until the next call.
	int run(const ScanoutPlan &plan, const VSyncInfo &info, uint8_t *vram, uint32_t sample_log2,
	        hipStream_t stream, ScanoutResult &result);
	void reset_history();
	// Accumulated device time of the scanout launches, when timed (milliseconds).
	void set_timing(bool enable) { timing_ = enable; }
	double poll_ms(bool wait);

private:
	struct Image
	{
		uint32_t *ptr = nullptr;
		size_t cap = 0; // pixels
	};
	Image circuit_[2];
	Image pool_[4];      // merged frames; the deinterlacer's field history points into it
	Image deinterlaced_;
	int history_[4] = { -1, -1, -1, -1 }; // pool index of field 0 (newest) .. 3, or -1
	uint32_t history_w_ = 0, history_h_ = 0;

	bool timing_ = false;
	struct Timed { hipEvent_t start, stop; };
	std::vector<Timed> timed_;

	int ensure(Image &img, size_t pixels);
};
}
