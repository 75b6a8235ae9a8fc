// Host front end: GIF parsing, register file and transfer state machine.
#include "gs_interface.hpp"
#include <cstring>
#include <algorithm>
#include <chrono>

namespace pgs
{
namespace
{
// log2 of a sample count 1/2/4/8/16, or -1 for any other value.
int sample_count_log2(uint32_t s)
{
	switch (s)
	{
	case 1: return 0;
	case 2: return 1;
	case 4: return 2;
	case 8: return 3;
	case 16: return 4;
	default: return -1;
	}
}
}

bool GSInterface::init(const GSOptions &options)
{
	options_ = options;
	const int l = sample_count_log2(uint32_t(options.super_sampling));
	if (l < 0)
		return false;
	const uint32_t sample_log2 = uint32_t(l);
	// Sampling grids (gs/gs_interface.cpp:65-117): 2x = 1x2, 4x = 2x2 or sparse 1x4, 8x = 2x4,
	// 16x = 4x4 or sparse 2x8. Super-sampled textures are off at 1x (gs/gs_interface.cpp:74-78).
	rp_.super_sampled_textures = options.super_sampled_textures && sample_log2 != 0;
	const bool sparse = !options.ordered_super_sampling && (sample_log2 == 2 || sample_log2 == 4);
	renderer_.set_sparse_grid(sparse);
	const uint32_t capacity_log2 = options.dynamic_super_sampling ? 4u : sample_log2;
	if (!renderer_.init(sample_log2, options.upload_arena_bytes, options.sampler_feedback, options.texture_stream,
	                    capacity_log2))
		return false;
	rp_.init(&renderer_, sample_log2);
	rp_.set_sample_log2(sample_log2, sparse);
	rp_.host = this;
	reset_registers();
	return true;
}

// gs/gs_interface.cpp:26-39: pending work is submitted first (a host->local transfer keeps
// what arrived and stays active, as the reference's flush leaves it), then the register file,
// GIF paths, vertex queue and cached CLUT base pointers are reset.
void GSInterface::reset_context_state()
{
	flush();
	reset_registers();
}

void GSInterface::reset_registers()
{
	regs_ = {};
	for (auto &p : paths_)
		p = {};
	cached_cbp_[0] = cached_cbp_[1] = 0;
	rp_.reset_vertex_queue();
	rp_.mark_state_dirty();
}

void GSInterface::clobber_register_state()
{
	flush();
	rp_.reset_vertex_queue();
}

bool GSInterface::set_super_sampling_rate(SuperSampling rate, bool ordered_grid, bool super_sampled_textures)
{
	const int l = sample_count_log2(uint32_t(rate));
	if (l < 0)
		return false;
	const uint32_t sample_log2 = uint32_t(l);
	if (sample_log2 > renderer_.capacity_log2())
		return false;
	flush();
	// At 1x the reference turns super-sampled textures off (gs/gs_interface.cpp:74-78).
	rp_.super_sampled_textures = super_sampled_textures && sample_log2 != 0;
	options_.super_sampled_textures = rp_.super_sampled_textures;
	const bool sparse = !ordered_grid && (sample_log2 == 2 || sample_log2 == 4);
	renderer_.set_sparse_grid(sparse);
	if (!renderer_.set_sample_log2(sample_log2))
		return false;
	rp_.set_sample_log2(sample_log2, sparse);
	options_.super_sampling = rate;
	options_.ordered_super_sampling = ordered_grid;
	return true;
}

// Kernel kinds of the renderer's launch timers (Renderer::LaunchKind) by TimestampType.
void GSInterface::poll_timestamps(bool wait)
{
	double ms[16] = {};
	renderer_.poll_timings(ms, 16, wait);
	auto add = [&](TimestampType t, double v) { timestamp_ms_[uint32_t(t)] += v; };
	add(TimestampType::CopyVRAM, ms[0] + ms[1]);
	add(TimestampType::TriangleSetup, ms[2] + ms[7]);
	add(TimestampType::Binning, ms[3]);
	add(TimestampType::Shading, ms[4] + ms[8]);
	add(TimestampType::PaletteUpdate, ms[5]);
	add(TimestampType::TextureUpload, ms[6] + ms[9]);
	add(TimestampType::VSync, scanout_.poll_ms(wait));
}

void GSInterface::set_timestamps(bool enable)
{
	if (timestamps_ && !enable)
		poll_timestamps(true);
	timestamps_ = enable;
	renderer_.profile_enable(enable);
	scanout_.set_timing(enable);
}

int GSInterface::vsync(const VSyncInfo &info_in, ScanoutResult &result)
{
	result = ScanoutResult{};
	const CrtcRegisters r = decode_crtc_registers(priv_.data());
	uint32_t xl, yl;
	renderer_.rates(xl, yl);
	// The reference switches its renderer to field-aware super-sampling here
	// (gs/gs_interface.cpp:4666-4672, gs/gs_renderer.cpp:1953, 2123, 2605); not built.
	if (info_in.high_resolution_scanout && yl && r.ffmd && r.interlace && r.cmod != 0)
		return kVsyncUnsupported;
	// A read circuit shows PSMCT32/24/16/16S; the reference reads other formats' addresses as
	// word indices (sample_circuit.frag:50-71), which this backend does not reproduce.
	auto displayable = [](uint32_t psm) {
		return psm == PSMCT32 || psm == PSMCT24 || psm == PSMCT16 || psm == PSMCT16S;
	};
	for (uint32_t c = 0; c < 2; c++)
		if (r.en[c] && !displayable(r.dispfb[c].psm))
			return kVsyncUnsupported;
	if (renderer_.recording())
		return kVsyncUnsupported;

	VSyncInfo info = info_in;
	if (rp_.complex_scanmsk_timeout)
	{
		info.force_progressive = false;
		rp_.complex_scanmsk_timeout--;
	}
	flush();
	const ScanoutPlan plan = plan_scanout(r, info, xl, yl);
	hipStream_t stream = renderer_.begin_external_work();
	const int rc = scanout_.run(plan, info, renderer_.vram_device(), renderer_.sample_log2(), stream, result);
	renderer_.end_external_work(plan.extwrite);
	if (plan.extwrite)
		rp_.writes().write(page_span(plan.exbp, plan.exbw, PSMCT32, plan.wdx, plan.wdy, plan.wdx + plan.ext_width - 1u,
		                             plan.wdy + plan.ext_height - 1u));
	if (timestamps_)
		poll_timestamps(false);
	return rc;
}

bool GSInterface::vsync_can_skip(const VSyncInfo &info) const
{
	return rp_.complex_scanmsk_timeout == 0 && !scanout_is_interlaced(decode_crtc_registers(priv_.data()), info);
}

double GSInterface::get_accumulated_timestamps(TimestampType type)
{
	if (uint32_t(type) >= uint32_t(TimestampType::Count))
		return -1.0;
	if (timestamps_)
		poll_timestamps(true);
	return timestamp_ms_[uint32_t(type)] * 1e-3;
}

// Transfers and draws execute in submission order on one stream. The host keeps at most one
// kind of work pending: queuing a transfer first flushes pending primitives, and a draw
// first flushes pending transfers (see RenderPassBuilder). The reference reaches the same
// ordering through page-hazard flushes (page_tracker.cpp:221-325, 621-673).
void GSInterface::begin_transfer_work()
{
	if (rp_.has_primitives())
		flush_render_pass();
}

void GSInterface::flush_render_pass()
{
	rp_.flush();
}

void GSInterface::before_first_primitive()
{
	renderer_.flush_transfers();
}

void GSInterface::before_texture_read()
{
	if (xfer_.host_to_local_active)
		flush_pending_transfer(true);
}

// Registers that feed primitive state mark the state dirty when their value changes
// (update_internal_register, gs/gs_interface.cpp:3877-3885); a dirty state blocks
// parallelogram fusion of the next triangle.
void GSInterface::set_state_reg(uint64_t &reg, uint64_t v)
{
	if (reg != v)
	{
		reg = v;
		rp_.mark_state_dirty();
	}
}

// ---------------------------------------------------------------------------------------
// GIF packet decoding (gs/gs_interface.cpp:4290-4397).

void GSInterface::gif_transfer(uint32_t path_index, const void *data, size_t size)
{
	auto &path = paths_[path_index & 3u];
	size /= 16;
	const uint64_t *q = static_cast<const uint64_t *>(data);

	for (size_t i = 0; i < size;)
	{
		if (path.loop == path.nloop)
		{
			uint64_t lo = q[2 * i], hi = q[2 * i + 1];
			path.tag_lo = lo;
			path.nloop = uint32_t(lo & 0x7fff);
			path.flg = uint32_t((lo >> 58) & 3u);
			path.nreg = uint32_t((lo >> 60) & 0xf);
			path.regs = hi;
			path.loop = 0;
			path.reg = 0;
			bool pre = ((lo >> 46) & 1u) != 0;
			if (path.flg == 0 && pre && path.nloop)
				write_register(0x00, (lo >> 47) & 0x7ffu);
			if (path.nloop)
				regs_.internal_q = 1.0f;
			i++;
			continue;
		}

		const uint32_t nreg = path.nreg ? path.nreg : 16u;
		if (path.flg == 0)
		{
			// PACKED: one register per qword.
			while (i < size && path.loop < path.nloop)
			{
				uint32_t addr = uint32_t(path.regs >> (4u * path.reg)) & 0xfu;
				packed_write(addr, q + 2 * i);
				i++;
				if (++path.reg == nreg)
				{
					path.reg = 0;
					path.loop++;
				}
			}
		}
		else if (path.flg == 1)
		{
			// REGLIST: two 64-bit register values per qword.
			for (uint32_t j = 0; j < 2; j++)
			{
				uint32_t addr = uint32_t(path.regs >> (4u * path.reg)) & 0xfu;
				if (addr < 0xe)
					write_register(addr, q[2 * i + j]);
				if (++path.reg == nreg)
				{
					path.reg = 0;
					path.loop++;
					if (path.loop == path.nloop)
						break;
				}
			}
			i++;
		}
		else
		{
			// IMAGE: raw HWREG payload.
			size_t n = std::min<size_t>(size - i, path.nloop - path.loop);
			hwreg(q + 2 * i, n * 2);
			i += n;
			path.loop += uint32_t(n);
		}
	}
}

// PACKED register formats (gs/gs_interface.cpp:3992-4138; gs/gs_registers.hpp:443-517).
void GSInterface::packed_write(uint32_t addr, const uint64_t *w)
{
	const uint64_t lo = w[0], hi = w[1];
	switch (addr)
	{
	case 0x0:
		write_register(0x00, lo & 0x7ffu);
		break;
	case 0x1:
	{
		uint64_t rgba = (lo & 0xff) | (((lo >> 32) & 0xff) << 8) | ((hi & 0xff) << 16) | (((hi >> 32) & 0xff) << 24);
		uint32_t qbits;
		memcpy(&qbits, &regs_.internal_q, 4);
		regs_.rgbaq = rgba | (uint64_t(qbits) << 32);
		break;
	}
	case 0x2:
		regs_.st = lo;
		{
			uint32_t qbits = uint32_t(hi);
			memcpy(&regs_.internal_q, &qbits, 4);
		}
		break;
	case 0x3:
		regs_.uv = (lo & 0x3fffu) | (((lo >> 32) & 0x3fffu) << 16);
		break;
	case 0x4: case 0xc:
	{
		uint64_t x = lo & 0xffff, y = (lo >> 32) & 0xffff;
		uint64_t z = (hi >> 4) & 0xffffff, f = (hi >> 36) & 0xff;
		bool adc = addr == 0xc || ((hi >> 47) & 1u);
		if (!adc)
			write_register(0x04, x | y << 16 | z << 32 | f << 56);
		else
			write_register(0x0c, x | y << 16 | z << 32 | f << 56);
		break;
	}
	case 0x5: case 0xd:
	{
		uint64_t x = lo & 0xffff, y = (lo >> 32) & 0xffff;
		uint64_t z = hi & 0xffffffffu;
		bool adc = addr == 0xd || ((hi >> 47) & 1u);
		write_register(adc ? 0x0d : 0x05, x | y << 16 | z << 32);
		break;
	}
	case 0x6: case 0x7: case 0x8: case 0x9:
		write_register(addr, lo);
		break;
	case 0xa:
		regs_.fog = ((hi >> 36) & 0xff) << 56;
		break;
	case 0xe:
		write_register(uint32_t(hi & 0x7f), lo);
		break;
	default:
		break;
	}
}

// A+D register writes (gs/gs_interface.cpp:3507-3986).
void GSInterface::write_register(uint32_t addr, uint64_t v)
{
	switch (addr)
	{
	case 0x00: write_prim(v); break;
	case 0x01: case 0x11: regs_.rgbaq = v; break;
	case 0x02: regs_.st = v; break;
	case 0x03: regs_.uv = v; break;
	case 0x0a: regs_.fog = v; break;
	case 0x04: case 0x0c:
		rp_.vertex_kick(regs_, uint32_t(v & 0xffff), uint32_t((v >> 16) & 0xffff), uint32_t((v >> 32) & 0xffffff),
		                float(uint32_t(v >> 56)), addr == 0x0c);
		break;
	case 0x05: case 0x0d:
		rp_.vertex_kick(regs_, uint32_t(v & 0xffff), uint32_t((v >> 16) & 0xffff), uint32_t(v >> 32),
		                float(uint32_t(regs_.fog >> 56)), addr == 0x0d);
		break;
	case 0x06: case 0x07: write_tex0(addr - 0x06, v); break;
	case 0x08: case 0x09: set_state_reg(regs_.clamp[addr - 0x08], v); break;
	case 0x14: case 0x15: set_state_reg(regs_.tex1[addr - 0x14], v); break;
	case 0x16: case 0x17:
	{
		// TEX2 rewrites the palette / PSM part of TEX0 (gs/gs_interface.cpp:2950-2968).
		const uint64_t keep = 0x3fffull | (0x3full << 14) | (0xfull << 26) | (0xfull << 30) | (1ull << 34) | (3ull << 35);
		uint32_t ctx = addr - 0x16;
		write_tex0(ctx, (regs_.tex0[ctx] & keep) | (v & ~keep));
		break;
	}
	case 0x18: case 0x19: regs_.xyoffset[addr - 0x18] = v; break;
	case 0x1a: regs_.prmodecont = v; break;
	case 0x1b: write_prmode(v); break;
	case 0x1c: regs_.texclut = v; break;
	case 0x22: set_state_reg(regs_.scanmsk, v); break;
	case 0x34: case 0x35: set_state_reg(regs_.miptbp1[addr - 0x34], v); break;
	case 0x36: case 0x37: set_state_reg(regs_.miptbp2[addr - 0x36], v); break;
	case 0x3b: set_state_reg(regs_.texa, v); break;
	case 0x3d: regs_.fogcol = v; break;
	// SCISSOR does not block parallelogram fusion: the reference clears its dirty bit when the
	// next primitive is clipped, before the fusion test (gs/gs_interface.cpp:2546, 2584).
	// A scissor change ends parallelogram fusion like any state change (STATE_DIRTY_SCISSOR_BIT,
	// gs/gs_interface.cpp:2546, 2584).
	case 0x40: case 0x41: set_state_reg(regs_.scissor[addr - 0x40], v); break;
	case 0x42: case 0x43: set_state_reg(regs_.alpha[addr - 0x42], v); break;
	case 0x44: set_state_reg(regs_.dimx, v); break;
	case 0x45: set_state_reg(regs_.dthe, v); break;
	case 0x46: set_state_reg(regs_.colclamp, v); break;
	case 0x47: case 0x48: set_state_reg(regs_.test[addr - 0x47], v); break;
	case 0x49: set_state_reg(regs_.pabe, v); break;
	case 0x4a: case 0x4b: set_state_reg(regs_.fba[addr - 0x4a], v); break;
	case 0x4c: case 0x4d: set_state_reg(regs_.frame[addr - 0x4c], v); break;
	case 0x4e: case 0x4f: set_state_reg(regs_.zbuf[addr - 0x4e], v); break;
	case 0x50: regs_.bitbltbuf = v; break;
	case 0x51: regs_.trxpos = v; break;
	case 0x52: regs_.trxreg = v; break;
	case 0x53: regs_.trxdir = v; init_transfer(); break;
	case 0x54: hwreg(&v, 1); break;
	// SIGNAL / FINISH / LABEL (gs/gs_interface.cpp:3970-3986): the caller's handler decides
	// whether pending work is submitted.
	case 0x60: if (signal_.on_signal && signal_.on_signal(signal_.user, v)) flush(); break;
	case 0x61: if (signal_.on_finish && signal_.on_finish(signal_.user, v)) flush(); break;
	case 0x62: if (signal_.on_label && signal_.on_label(signal_.user, v)) flush(); break;
	default: break;
	}
}

// PRIM (gs/gs_interface.cpp:3507-3556). With PRMODECONT.AC = 1 the whole register is written;
// with AC = 0 only the primitive type changes and the other bits keep what PRMODE set. A type
// change marks the state dirty only when the written AA1 bit is set (PRIM_TEMPLATE there).
void GSInterface::write_prim(uint64_t v)
{
	v &= 0x7ffu;
	if (regs_.prmodecont & 1u)
		set_state_reg(regs_.prim, v);
	else
	{
		if ((regs_.prim & 7u) != (v & 7u) && ((v >> 7) & 1u))
			rp_.mark_state_dirty();
		regs_.prim = (regs_.prim & ~7ull) | (v & 7u);
	}
	rp_.reset_vertex_queue();
}

// PRMODE (gs/gs_interface.cpp:3692-3722): with PRMODECONT.AC = 0 it supplies every PRIM bit but
// the primitive type; with AC = 1 the write is ignored.
void GSInterface::write_prmode(uint64_t v)
{
	regs_.prmode = v;
	if (regs_.prmodecont & 1u)
		return;
	set_state_reg(regs_.prim, (v & 0x7f8u) | (regs_.prim & 7u));
}

// TEX0 write: texture state change plus the CLUT load its CLD field asks for.
void GSInterface::write_tex0(uint32_t ctx, uint64_t v)
{
	set_state_reg(regs_.tex0[ctx], v);
	handle_clut_upload(ctx);
}

// CLUT load decision and parameters (gs/gs_interface.cpp:550-797): CLD 1 always loads, 2/3
// load and record CBP0/CBP1, 4/5 load only when CBP differs from CBP0/CBP1. The load reads
// VRAM as of this register write, so a partially received upload, pending transfers and
// pending primitives are all executed first.
void GSInterface::handle_clut_upload(uint32_t ctx)
{
	const uint64_t t0 = regs_.tex0[ctx];
	const uint32_t psm = uint32_t((t0 >> 20) & 0x3f);
	if (psm != PSMT8 && psm != PSMT4 && psm != PSMT8H && psm != PSMT4HL && psm != PSMT4HH)
		return;
	const uint32_t cld = uint32_t((t0 >> 61) & 7);
	const uint32_t cbp = uint32_t((t0 >> 37) & 0x3fff);
	bool load = false;
	switch (cld)
	{
	case 1: load = true; break;
	case 2: case 3: load = true; cached_cbp_[cld & 1] = cbp; break;
	case 4: case 5: load = cached_cbp_[cld & 1] != cbp; cached_cbp_[cld & 1] = cbp; break;
	default: break;
	}
	if (!load)
		return;
	if (xfer_.host_to_local_active)
		flush_pending_transfer(true);
	flush_render_pass();
	renderer_.flush_transfers();

	pgs_clut_args a{};
	uint32_t cpsm = uint32_t((t0 >> 51) & 0xf);
	if (cpsm == PSMCT24)
		cpsm = PSMCT32;
	a.cpsm = cpsm;
	a.cbp = cbp;
	a.csm = uint32_t((t0 >> 55) & 1);
	a.csa = uint32_t((t0 >> 56) & 0x1f);
	if (cpsm == PSMCT32)
		a.csa &= 15u; // a 32-bit CLUT only targets the lower banks (gs/gs_interface.cpp:608-610)
	a.index4 = (psm == PSMT4 || psm == PSMT4HL || psm == PSMT4HH) ? 1u : 0u;
	if (a.csm)
	{
		a.cbw = uint32_t(regs_.texclut & 0x3f);
		a.cou = uint32_t((regs_.texclut >> 6) & 0x3f);
		a.cov = uint32_t((regs_.texclut >> 12) & 0x3ff);
	}
	rp_.writes().clut_loaded();
	renderer_.clut_load(a);
}

// ---------------------------------------------------------------------------------------
// Transfers (gs/gs_interface.cpp:3168-3276, 2970-3148).

static CopyDesc decode_copy(const RegisterState &r)
{
	CopyDesc c{};
	uint64_t b = r.bitbltbuf, p = r.trxpos, g = r.trxreg;
	c.sbp = uint32_t(b & 0x3fff);
	c.sbw = uint32_t((b >> 16) & 0x3f);
	c.spsm = uint32_t((b >> 24) & 0x3f);
	c.dbp = uint32_t((b >> 32) & 0x3fff);
	c.dbw = uint32_t((b >> 48) & 0x3f);
	c.dpsm = uint32_t((b >> 56) & 0x3f);
	c.ssax = uint32_t(p & 0x7ff);
	c.ssay = uint32_t((p >> 16) & 0x7ff);
	c.dsax = uint32_t((p >> 32) & 0x7ff);
	c.dsay = uint32_t((p >> 48) & 0x7ff);
	c.w = uint32_t(g & 0xfff);
	c.h = uint32_t((g >> 32) & 0xfff);
	return c;
}

void GSInterface::init_transfer()
{
	flush_pending_transfer(false);
	CopyDesc c = decode_copy(regs_);
	uint32_t xdir = uint32_t(regs_.trxdir & 3u);

	if (xdir == 2)
	{
		if (c.w == 0 || c.h == 0)
			return;
		begin_transfer_work();
		rp_.writes().write(page_span(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u));
		rp_.writes().mark_super_sample(
		    page_rect(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u), false);
		renderer_.queue_local_copy(c);
	}
	else if (xdir == 0)
	{
		// The texture cache must not reuse an image of the destination from here on: a texture
		// lookup can come before the payload is flushed (each flush logs its pages again).
		rp_.writes().write(page_span(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u));
		if (c.w && c.h)
			rp_.writes().mark_super_sample(
			    page_rect(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u), false);
		xfer_.copy = c;
		xfer_.required_qwords = uint32_t((uint64_t(c.w) * c.h * psm_bpp(c.dpsm)) / 64u);
		xfer_.host_to_local_active = xfer_.required_qwords != 0;
		xfer_.payload.clear();
		xfer_.flushed_pixels = 0;
	}
	else if (xdir == 1)
	{
		// Local->host readback into the transfer FIFO (gs/gs_interface.cpp:3213-3275). The read
		// waits for every earlier draw and transfer, as the reference's host-read timeline does;
		// 4-bit formats are not read back (gs/gs_util.hpp:49-110) and leave the FIFO zero.
		uint32_t bytes = uint32_t((uint64_t(c.w) * c.h * psm_bpp(c.spsm)) / 8u);
		xfer_.fifo.assign(bytes, 0);
		xfer_.fifo_offset_128 = 0;
		xfer_.fifo_size_128 = bytes / 16u;
		// Hacks::unsynced_readbacks: zeros instead of waiting for unfinished GPU work.
		if (hacks_.unsynced_readbacks && (rp_.has_primitives() || renderer_.has_pending_transfers() || !renderer_.idle()))
			return;
		const auto t0 = std::chrono::steady_clock::now();
		flush_render_pass();
		renderer_.flush_transfers();
		if (host_vram_.size() < kVramSize)
			host_vram_.resize(kVramSize);
		renderer_.read_vram(host_vram_.data(), 0, 1);
		if (timestamps_)
			timestamp_ms_[uint32_t(TimestampType::Readback)] +=
			    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		const uint8_t *v = host_vram_.data();
		for (uint32_t y = 0; y < c.h; y++)
			for (uint32_t x = 0; x < c.w; x++)
			{
				uint32_t idx = y * c.w + x;
				uint32_t a = swizzle_addr((x + c.ssax) & 2047u, (y + c.ssay) & 2047u, c.sbp, c.sbw, c.spsm);
				switch (c.spsm)
				{
				case PSMCT32: case PSMZ32:
					memcpy(&xfer_.fifo[idx * 4], v + a, 4);
					break;
				case PSMCT24: case PSMZ24:
					if (idx * 3 + 3 <= bytes)
						memcpy(&xfer_.fifo[idx * 3], v + a, 3);
					break;
				case PSMCT16: case PSMCT16S: case PSMZ16: case PSMZ16S:
					memcpy(&xfer_.fifo[idx * 2], v + a, 2);
					break;
				case PSMT8:
					xfer_.fifo[idx] = v[a];
					break;
				case PSMT8H:
					xfer_.fifo[idx] = v[a | 3u];
					break;
				default:
					break;
				}
			}
	}
}

void GSInterface::hwreg(const uint64_t *data, size_t count)
{
	if (!xfer_.host_to_local_active)
		return;
	xfer_.payload.insert(xfer_.payload.end(), data, data + count);
	if (xfer_.payload.size() >= xfer_.required_qwords)
		flush_pending_transfer(false);
}

// Emits the pixels whose payload has arrived. With keep_alive the transfer stays active
// and later payload continues where this flush ended (partial transfers).
void GSInterface::flush_pending_transfer(bool keep_alive)
{
	if (xfer_.host_to_local_active)
	{
		const CopyDesc &c = xfer_.copy;
		uint32_t total = c.w * c.h;
		uint64_t bits = uint64_t(xfer_.payload.size()) * 64u;
		uint32_t avail = uint32_t(std::min<uint64_t>(total, bits / psm_bpp(c.dpsm)));
		if (avail > xfer_.flushed_pixels)
		{
			begin_transfer_work();
			CopyDesc d = c;
			d.first_pixel = xfer_.flushed_pixels;
			d.end_pixel = avail;
			rp_.writes().write(page_span(c.dbp, c.dbw, c.dpsm, c.dsax, c.dsay, c.dsax + c.w - 1u, c.dsay + c.h - 1u));
			renderer_.queue_host_copy(d, xfer_.payload.data(), xfer_.payload.size() * 8u);
			xfer_.flushed_pixels = avail;
		}
	}

	if (!keep_alive)
	{
		xfer_.host_to_local_active = false;
		xfer_.payload.clear();
		xfer_.flushed_pixels = 0;
	}
}

// gs/gs_interface.cpp:3150-3166
void GSInterface::read_transfer_fifo(void *data, uint32_t n)
{
	uint32_t avail = xfer_.fifo_size_128 - std::min(xfer_.fifo_size_128, xfer_.fifo_offset_128);
	uint32_t copy = std::min(n, avail);
	if (copy)
		memcpy(data, xfer_.fifo.data() + 16u * xfer_.fifo_offset_128, copy * 16u);
	if (copy < n)
		memset(static_cast<uint8_t *>(data) + 16u * copy, 0, (n - copy) * 16u);
	xfer_.fifo_offset_128 += copy;
}

// ---------------------------------------------------------------------------------------
// Flush and host VRAM access (gs/gs_interface.cpp:4184-4264).

void GSInterface::flush()
{
	flush_pending_transfer(true);
	flush_render_pass();
	renderer_.flush_transfers();
	if (timestamps_)
		poll_timestamps(false);
}

const void *GSInterface::map_vram_read(size_t offset, size_t size)
{
	flush();
	if (host_vram_.size() < kVramSize)
		host_vram_.resize(kVramSize);
	renderer_.read_vram(host_vram_.data(), 0, 1);
	if (offset + size > kVramSize)
		return nullptr;
	return host_vram_.data() + offset;
}

void GSInterface::read_all_planes(void *dst)
{
	flush();
	renderer_.read_vram(dst, 0, renderer_.num_planes());
}

void *GSInterface::map_vram_write(size_t offset, size_t size)
{
	if (offset + size > kVramSize)
		return nullptr;
	map_vram_read(0, kVramSize);
	return host_vram_.data() + offset;
}

void GSInterface::end_vram_write(size_t offset, size_t size)
{
	if (offset + size > kVramSize || size == 0)
		return;
	// The written 8 KiB pages go back to the device whole, as the reference commits a host write
	// with every block of its pages marked (gs/gs_interface.cpp:4210-4228, gs/gs_renderer.cpp:
	// 1478-1539); the texture cache forgets images decoded from them.
	const uint32_t first_page = uint32_t(offset / kPageBytes), last_page = uint32_t((offset + size - 1u) / kPageBytes);
	rp_.writes().write(PageSpan{ first_page, last_page });
	rp_.writes().mark_super_sample(PageRect{ first_page, 0u, last_page - first_page + 1u, 1u }, false);
	const auto t0 = std::chrono::steady_clock::now();
	renderer_.write_vram(host_vram_.data() + size_t(first_page) * kPageBytes, size_t(first_page) * kPageBytes,
	                     size_t(last_page - first_page + 1u) * kPageBytes);
	if (timestamps_)
		timestamp_ms_[uint32_t(TimestampType::SyncHostToVRAM)] +=
		    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}
