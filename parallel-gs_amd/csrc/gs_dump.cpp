// GS dump reader and writer (see gs_dump.hpp).
#include "gs_dump.hpp"
#include <algorithm>
#include <cstring>

namespace pgs
{
namespace
{
// DumpHeader (dump/gs_dump_parser.hpp:25-36): nine 32-bit words.
struct DumpHeader
{
	uint32_t version, state_size, serial_offset, serial_size, crc;
	uint32_t screenshot_width, screenshot_height, screenshot_offset, screenshot_size;
};
static_assert(sizeof(DumpHeader) == 36, "dump header layout");

constexpr uint32_t kTexflush = 0x3f;
constexpr uint32_t kGifAddrAD = 0xe;

uint64_t gif_tag_lo(uint32_t nloop, bool eop, bool pre, uint32_t prim, uint32_t flg, uint32_t nreg)
{
	return uint64_t(nloop & 0x7fffu) | (uint64_t(eop) << 15) | (uint64_t(pre) << 46) |
	       (uint64_t(prim & 0x7ffu) << 47) | (uint64_t(flg & 3u) << 58) | (uint64_t(nreg & 15u) << 60);
}
} // namespace

// ---------------------------------------------------------------------------------------
// Reader (dump/gs_dump_parser.cpp).

void DumpReader::read(void *dst, size_t n)
{
	if (fread(dst, 1, n, file_.get()) != n)
		eof_ = true;
}

uint8_t DumpReader::u8()
{
	uint8_t v = 0;
	read(&v, 1);
	return v;
}

uint32_t DumpReader::u32()
{
	uint32_t v = 0;
	read(&v, 4);
	return v;
}

uint64_t DumpReader::u64()
{
	uint64_t v = 0;
	read(&v, 8);
	return v;
}

void DumpReader::skip(size_t n)
{
	if (fseek(file_.get(), long(n), SEEK_CUR) != 0)
		eof_ = true;
}

bool DumpReader::open(const char *path, GSInterface &iface, bool raw)
{
	iface_ = &iface;
	raw_ = raw;
	file_.reset(fopen(path, "rb"));
	if (!file_)
		return false;
	return raw ? true : restart();
}

// The register state block (gs_dump_parser.cpp:86-135): registers are assigned without their
// write side effects; UV and FOG carry their low 32 bits only.
void DumpReader::read_register_state()
{
	RegisterState &r = iface_->get_register_state();
	const uint32_t version = u32();
	r.prim = u64();
	r.prmodecont = u64();
	r.texclut = u64();
	r.scanmsk = u64();
	r.texa = u64();
	r.fogcol = u64();
	r.dimx = u64();
	r.dthe = u64();
	r.colclamp = u64();
	r.pabe = u64();
	r.bitbltbuf = u64();
	r.trxdir = u64();
	r.trxpos = u64();
	r.trxreg = u64();
	(void)u64();
	for (uint32_t c = 0; c < 2; c++)
	{
		r.xyoffset[c] = u64();
		r.tex0[c] = u64();
		r.tex1[c] = u64();
		r.clamp[c] = u64();
		r.miptbp1[c] = u64();
		r.miptbp2[c] = u64();
		r.scissor[c] = u64();
		r.alpha[c] = u64();
		r.test[c] = u64();
		r.fba[c] = u64();
		r.frame[c] = u64();
		r.zbuf[c] = u64();
	}
	r.rgbaq = u64();
	r.st = u64();
	r.uv = (r.uv & ~0xffffffffull) | u32();
	r.fog = (r.fog & ~0xffffffffull) | u32();
	(void)u64(); // XYZ
	(void)u32(); // GIFReg
	(void)u32();
	(void)u32(); // transfer X / Y
	(void)u32();
	if (version >= 9)
		skip(61);
}

bool DumpReader::restart()
{
	if (!file_)
		return false;
	rewind(file_.get());
	eof_ = false;
	if (raw_)
		return true;
	if (u32() != 0xffffffffu)
		return false;
	const uint32_t header_size = u32();
	if (header_size < sizeof(DumpHeader))
		return false;
	DumpHeader h{};
	read(&h, sizeof(h));
	if (h.version < kDumpStateVersionMin || h.version > kDumpStateVersion)
		return false;
	skip(header_size - sizeof(DumpHeader)); // serial and screenshot

	read_register_state();
	void *vram = iface_->map_vram_write(0, kVramSize);
	read(vram, kVramSize);
	iface_->end_vram_write(0, kVramSize);
	iface_->write_register(kTexflush, 0);

	for (uint32_t i = 0; i < 4; i++)
	{
		GIFPath &p = iface_->get_gif_path(i);
		uint64_t tag[2];
		read(tag, sizeof(tag));
		p.tag_lo = tag[0];
		p.nloop = uint32_t(tag[0] & 0x7fffu);
		p.flg = uint32_t((tag[0] >> 58) & 3u);
		p.nreg = uint32_t((tag[0] >> 60) & 15u);
		p.regs = tag[1];
		p.loop = 0; // the format stores the loops left as NLOOP
		p.reg = u32();
	}
	float q;
	read(&q, 4);
	iface_->get_register_state().internal_q = q;
	read(iface_->get_priv_register_state(), 8192);
	iface_->clobber_register_state();
	return !eof_;
}

bool DumpReader::iterate_until_vsync(bool high_res_scanout, bool conservative_crtc)
{
	if (!file_)
		return false;
	bool has_transfer = false;
	while (!eof_)
	{
		const DumpPacket type = DumpPacket(u8());
		if (eof_)
			break;
		switch (type)
		{
		case DumpPacket::Transfer:
		{
			const uint32_t path = u8();
			const uint32_t size = u32();
			buf_.resize(size);
			read(buf_.data(), size);
			if (!eof_)
				iface_->gif_transfer(path & 3u, buf_.data(), size & ~size_t(15));
			has_transfer = true;
			break;
		}
		case DumpPacket::Vsync:
		{
			// gs_dump_parser.cpp:164-186.
			VSyncInfo info;
			info.phase = u8();
			info.adapt_to_internal_horizontal_resolution = true;
			if (conservative_crtc)
			{
				info.skip_deinterlace = true;
				info.crtc_offsets = true;
				info.overscan = true;
			}
			else
			{
				info.high_resolution_scanout = high_res_scanout;
				info.force_progressive = true;
				info.anti_blur = true;
			}
			iface_->flush();
			vsync_rc_ = iface_->vsync(info, vsync_);
			if (has_transfer || conservative_crtc)
				return true;
			break;
		}
		case DumpPacket::PrivRegisters:
			read(iface_->get_priv_register_state(), 8192);
			break;
		case DumpPacket::ReadFIFO:
			// The reference does not replay readbacks either (gs_dump_parser.cpp:193-199).
			(void)u32();
			return true;
		default:
			break; // unknown packet bytes are skipped one at a time, as the reference does
		}
	}
	return !eof_;
}

// ---------------------------------------------------------------------------------------
// Writer (dump/gs_dump_generator.cpp).

void DumpWriter::put(const void *src, size_t n)
{
	if (ok_ && fwrite(src, 1, n, file_.get()) != n)
		ok_ = false;
}

bool DumpWriter::init(const char *path, GSInterface &iface)
{
	file_.reset(fopen(path, "wb"));
	ok_ = file_ != nullptr;
	if (!ok_)
		return false;
	const void *vram = iface.map_vram_read(0, kVramSize);

	u32(0xffffffffu); // fake CRC
	DumpHeader h{};
	h.version = kDumpStateVersion;
	h.state_size = kVramSize + (18 + 12 * 2) * 8 + 7 * 4 + 4 * 5 * 4 + 4;
	u32(sizeof(DumpHeader));
	put(&h, sizeof(h));

	// Register state, version 8 layout (gs_dump_generator.cpp:39-95).
	const RegisterState &r = iface.get_register_state();
	u32(kDumpStateVersionMin);
	for (uint64_t v : { r.prim, r.prmodecont, r.texclut, r.scanmsk, r.texa, r.fogcol, r.dimx, r.dthe, r.colclamp,
	                    r.pabe, r.bitbltbuf, r.trxdir, r.trxpos, r.trxreg, r.trxreg })
		u64(v);
	for (uint32_t c = 0; c < 2; c++)
		for (uint64_t v : { r.xyoffset[c], r.tex0[c], r.tex1[c], r.clamp[c], r.miptbp1[c], r.miptbp2[c],
		                    r.scissor[c], r.alpha[c], r.test[c], r.fba[c], r.frame[c], r.zbuf[c] })
			u64(v);
	u64(r.rgbaq);
	u64(r.st);
	u32(uint32_t(r.uv));
	u32(uint32_t(r.fog));
	u64(0);           // XYZ
	u32(0xffffffffu); // GIFReg
	u32(0xffffffffu);
	u32(0);           // transfer X / Y
	u32(0);

	put(vram, kVramSize);
	for (uint32_t i = 0; i < 4; i++)
	{
		const GIFPath &p = iface.get_gif_path(i);
		const uint64_t tag[2] = { (p.tag_lo & ~uint64_t(0x7fff)) | uint64_t((p.nloop - p.loop) & 0x7fffu), p.regs };
		put(tag, sizeof(tag));
		u32(p.reg);
	}
	const float q = r.internal_q;
	put(&q, 4);
	put(iface.get_priv_register_state(), 8192);
	return ok_;
}

void DumpWriter::transfer(uint32_t path, const void *qwords, size_t bytes)
{
	u8(uint8_t(DumpPacket::Transfer));
	u8(uint8_t(path));
	u32(uint32_t(bytes));
	put(qwords, bytes);
}

bool DumpWriter::write_register(uint32_t addr, uint64_t payload)
{
	if (!file_)
		return false;
	const uint64_t q[4] = { gif_tag_lo(1, true, false, 0, 0, 1), kGifAddrAD, payload, uint64_t(addr) };
	transfer(1, q, sizeof(q));
	return ok_;
}

bool DumpWriter::write_packed(uint64_t prim, const uint32_t *registers, uint32_t num_registers, uint32_t num_loops,
                              const void *data)
{
	if (!file_ || num_registers == 0 || num_registers > 16)
		return false;
	uint64_t regs = 0;
	for (uint32_t i = 0; i < num_registers; i++)
		regs |= uint64_t(registers[i] & 15u) << (4 * i);
	const size_t payload = size_t(16) * num_registers * num_loops;
	std::vector<uint8_t> q(16 + payload);
	const uint64_t tag[2] = { gif_tag_lo(num_loops, true, true, uint32_t(prim), 0, num_registers), regs };
	memcpy(q.data(), tag, 16);
	if (payload)
		memcpy(q.data() + 16, data, payload);
	transfer(1, q.data(), q.size());
	return ok_;
}

bool DumpWriter::write_vsync(uint32_t field, const GSInterface &iface)
{
	if (!file_)
		return false;
	u8(uint8_t(DumpPacket::PrivRegisters));
	put(iface.get_priv_register_state(), 8192);
	u8(uint8_t(DumpPacket::Vsync));
	u8(uint8_t(field));
	return ok_;
}

bool DumpWriter::write_image_upload(uint32_t addr, uint32_t psm, uint32_t width, uint32_t height, const void *data,
                                    size_t size)
{
	if (!file_ || size % 16)
		return false;
	const uint64_t bitbltbuf = (uint64_t(psm & 0x3fu) << 24) | (uint64_t((addr / 256u) & 0x3fffu) << 32) |
	                           (uint64_t((width / 64u) & 0x3fu) << 48) | (uint64_t(psm & 0x3fu) << 56);
	write_register(0x50, bitbltbuf);
	write_register(0x51, 0);
	write_register(0x52, uint64_t(width & 0xfffu) | (uint64_t(height & 0xfffu) << 32));
	write_register(0x53, 0);
	// IMAGE packets of at most 32767 qwords (NLOOP is a 15-bit field).
	const size_t max_bytes = size_t(0x7fff) * 16;
	const uint8_t *src = static_cast<const uint8_t *>(data);
	std::vector<uint8_t> q;
	for (size_t i = 0; i < size;)
	{
		const size_t n = std::min(size - i, max_bytes);
		const uint64_t tag[2] = { gif_tag_lo(uint32_t(n / 16), i + n == size, false, 0, 2, 0), 0 };
		q.resize(16 + n);
		memcpy(q.data(), tag, 16);
		memcpy(q.data() + 16, src + i, n);
		transfer(1, q.data(), q.size());
		i += n;
	}
	return ok_;
}

bool DumpWriter::close()
{
	if (!file_)
		return false;
	const bool ok = ok_ && fflush(file_.get()) == 0;
	file_.reset();
	return ok;
}
}
