// GS dump files: the PCSX2 GS-dump format (state versions 8 and 9) as a fixture, regression
// and save-state format for the backend.
//
// Reader: GSDumpParser (dump/gs_dump_parser.cpp:4-207) — a header, the register state, 4 MiB
// of VRAM, the four GIF paths, the internal Q and the privileged registers, then packets
// (Transfer, Vsync, ReadFIFO, PrivRegisters) replayed into a GSInterface. Writer:
// GSDumpGenerator (dump/gs_dump_generator.cpp:1-238) — the same start state taken from a live
// context, then register writes, PACKED GIF packets, image uploads and vsyncs.
#pragma once
#include "gs_interface.hpp"
#include <cstdio>
#include <memory>
#include <vector>

namespace pgs
{
// GSDumpPacketType (dump/gs_dump_parser.hpp:17-23).
enum class DumpPacket : uint8_t { Transfer = 0, Vsync = 1, ReadFIFO = 2, PrivRegisters = 3 };

constexpr uint32_t kDumpStateVersion = 9, kDumpStateVersionMin = 8;

class DumpReader
{
public:
	// open (header + start state) or open_raw (packets only, from the current state).
	bool open(const char *path, GSInterface &iface, bool raw);
	// Rewinds and loads the start state again (a save-state restore).
	bool restart();
	// Replays packets up to the next vsync after some GIF traffic (or every vsync with
	// conservative_crtc). False at the end of the file.
	bool iterate_until_vsync(bool high_res_scanout, bool conservative_crtc);
	// The last vsync's result, handed over once (GSDumpParser::consume_vsync_result).
	ScanoutResult consume_vsync_result()
	{
		ScanoutResult r = vsync_;
		vsync_ = ScanoutResult{};
		return r;
	}
	int vsync_status() const { return vsync_rc_; }

private:
	struct Closer { void operator()(FILE *f) const { if (f) fclose(f); } };
	std::unique_ptr<FILE, Closer> file_;
	GSInterface *iface_ = nullptr;
	bool raw_ = false;
	bool eof_ = false;
	std::vector<uint8_t> buf_;
	ScanoutResult vsync_{};
	int vsync_rc_ = 0;

	void read(void *dst, size_t n);
	uint8_t u8();
	uint32_t u32();
	uint64_t u64();
	void skip(size_t n);
	void read_register_state();
};

class DumpWriter
{
public:
	// Writes the header and the context's current state (GSDumpGenerator::init).
	bool init(const char *path, GSInterface &iface);
	// A+D write of one register as a one-qword PACKED transfer on path 1.
	bool write_register(uint32_t addr, uint64_t payload);
	// PACKED packet with PRE = 1: `num_registers` 4-bit register descriptors, `num_loops` loops of
	// 16-byte qwords from `data`.
	bool write_packed(uint64_t prim, const uint32_t *registers, uint32_t num_registers, uint32_t num_loops,
	                  const void *data);
	// Privileged registers, then a vsync of the given field.
	bool write_vsync(uint32_t field, const GSInterface &iface);
	// BITBLTBUF / TRXPOS / TRXREG / TRXDIR and IMAGE packets of at most 32768 qwords.
	bool write_image_upload(uint32_t addr, uint32_t psm, uint32_t width, uint32_t height, const void *data,
	                        size_t size);
	bool close();

private:
	struct Closer { void operator()(FILE *f) const { if (f) fclose(f); } };
	std::unique_ptr<FILE, Closer> file_;
	bool ok_ = true;
	void put(const void *src, size_t n);
	void u8(uint8_t v) { put(&v, 1); }
	void u32(uint32_t v) { put(&v, 4); }
	void u64(uint64_t v) { put(&v, 8); }
	void transfer(uint32_t path, const void *qwords, size_t bytes);
};
}
