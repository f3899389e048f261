// Presentation of the frame's RGBA8 target.  After every frame the reference copies its
// R8G8B8A8 storage image (volumetric/src/GaussianEngine.cpp:318) into the swapchain image
// (GaussianEngine::draw and recordTargetCopy, GaussianEngine.cpp:714-762, 865-875).  Here the
// same image is copied into a caller's pitched device surface on a stream (the memory of a
// swapchain or interop image), or written to a PNG file for callers without a window.
#include "tpd_ctx.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace tpd;

namespace {

// CRC-32 as the PNG specification defines it (ISO/IEC 15948, annex D: reflected polynomial
// 0xEDB88320, register preset to all ones and inverted at the end).
struct Crc32Table {
    uint32_t t[256];
    Crc32Table() {
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[n] = c;
        }
    }
};

uint32_t crc_update(uint32_t reg, const uint8_t* p, size_t n) {
    static const Crc32Table tab;
    for (size_t i = 0; i < n; ++i) reg = tab.t[(reg ^ p[i]) & 0xFFu] ^ (reg >> 8);
    return reg;
}

// Adler-32 of the zlib format (RFC 1950, section 8.2), 5552 bytes between reductions so that
// the 32-bit sums cannot overflow.
struct Adler32 {
    uint32_t a = 1, b = 0;
    void update(const uint8_t* p, size_t n) {
        while (n) {
            const size_t k = n < 5552 ? n : 5552;
            for (size_t i = 0; i < k; ++i) {
                a += p[i];
                b += a;
            }
            a %= 65521u;
            b %= 65521u;
            p += k;
            n -= k;
        }
    }
    uint32_t value() const { return (b << 16) | a; }
};

void put_be32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

// One PNG chunk: length, type, data, CRC over type and data.
bool write_chunk(std::FILE* f, const char type[4], const uint8_t* data, uint32_t len) {
    uint8_t head[8];
    put_be32(head, len);
    std::memcpy(head + 4, type, 4);
    uint32_t reg = crc_update(0xFFFFFFFFu, head + 4, 4);
    reg = crc_update(reg, data, len);
    uint8_t tail[4];
    put_be32(tail, ~reg);
    return std::fwrite(head, 1, 8, f) == 8 && (len == 0 || std::fwrite(data, 1, len, f) == len) &&
           std::fwrite(tail, 1, 4, f) == 4;
}

// Truecolour-with-alpha PNG, 8 bits per channel, no interlace, filter type 0 on every scanline.
// The zlib stream holds stored (uncompressed) deflate blocks of at most 65535 bytes, one block
// per IDAT chunk, so the file is the scanlines plus 11 bytes per 64 KiB: the writer never needs
// the whole image twice in memory and costs one pass over it.
int encode_png(const uint8_t* rgba, uint32_t w, uint32_t h, const char* path) {
    if (!rgba || !path) return fail(TPD_ERR_INVALID, "png: null argument");
    if (w == 0 || h == 0 || w > 0x7FFFFFFFu || h > 0x7FFFFFFFu)
        return fail(TPD_ERR_INVALID, "png: width and height must be 1 .. 2^31 - 1");
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(TPD_ERR_INVALID, std::string("png: cannot open ") + path);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    bool ok = std::fwrite(sig, 1, 8, f) == 8;
    uint8_t ihdr[13];
    put_be32(ihdr, w);
    put_be32(ihdr + 4, h);
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // colour type: RGBA
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering, type 0 used throughout
    ihdr[12] = 0;  // no interlace
    ok = ok && write_chunk(f, "IHDR", ihdr, 13);

    const uint64_t row = 1 + 4 * (uint64_t)w;  // filter byte + pixels
    const uint64_t total = row * h;
    constexpr uint32_t kBlock = 65535;
    std::vector<uint8_t> buf(2 + 5 + kBlock + 4);
    Adler32 adler;
    uint64_t pos = 0;
    bool first = true;
    while (ok && pos < total) {
        const uint32_t len = (uint32_t)std::min<uint64_t>(kBlock, total - pos);
        const bool last = pos + len == total;
        size_t o = 0;
        if (first) {
            buf[o++] = 0x78;  // CMF: deflate, 32 KiB window
            buf[o++] = 0x01;  // FLG: (0x78 * 256 + 0x01) % 31 == 0, no dictionary
            first = false;
        }
        buf[o++] = last ? 1 : 0;  // BFINAL, BTYPE = 00 (stored), padded to the byte boundary
        buf[o++] = (uint8_t)len;
        buf[o++] = (uint8_t)(len >> 8);
        buf[o++] = (uint8_t)~len;
        buf[o++] = (uint8_t)(~len >> 8);
        uint8_t* d = buf.data() + o;
        uint32_t filled = 0;
        while (filled < len) {
            const uint64_t y = (pos + filled) / row, x = (pos + filled) % row;
            if (x == 0) {
                d[filled++] = 0;  // filter type None
                continue;
            }
            const uint32_t n = (uint32_t)std::min<uint64_t>(row - x, len - filled);
            std::memcpy(d + filled, rgba + y * (row - 1) + (x - 1), n);
            filled += n;
        }
        adler.update(d, len);
        o += len;
        pos += len;
        if (last) {
            put_be32(buf.data() + o, adler.value());
            o += 4;
        }
        ok = write_chunk(f, "IDAT", buf.data(), (uint32_t)o);
    }
    ok = ok && write_chunk(f, "IEND", nullptr, 0);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TPD_OK : fail(TPD_ERR_INVALID, std::string("png: write failed: ") + path);
}

int save_device_rgba8(const uint8_t* d_rgba, uint32_t w, uint32_t h, hipStream_t wait_on, const char* path) {
    std::vector<uint8_t> host(4 * (size_t)w * h);
    hipError_t e = wait_on ? hipStreamSynchronize(wait_on) : hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(host.data(), d_rgba, host.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "png readback");
    return encode_png(host.data(), w, h, path);
}

int copy_pitched(const uint8_t* d_rgba, uint32_t w, uint32_t h, void* d_dst, uint64_t dst_pitch, hipStream_t s,
                 const char* what) {
    if (!d_dst) return fail(TPD_ERR_INVALID, std::string(what) + ": null surface");
    const size_t row = 4 * (size_t)w;
    if (dst_pitch == 0) dst_pitch = row;
    if (dst_pitch < row) return fail(TPD_ERR_INVALID, std::string(what) + ": pitch below 4 * width");
    hipError_t e = hipMemcpy2DAsync(d_dst, (size_t)dst_pitch, d_rgba, row, row, h, hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? TPD_OK : hip_fail(e, what);
}

}  // namespace

extern "C" {

int tpd_write_png(const uint8_t* rgba, uint32_t width, uint32_t height, const char* path) {
    return encode_png(rgba, width, height, path);
}

int tpd_save_png(tpd_ctx* c, const char* path) {
    if (!c || !path) return fail(TPD_ERR_INVALID, "tpd_save_png: bad argument");
    hipSetDevice(c->device);
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_save_png: nothing rendered");
    return save_device_rgba8(c->d_rgba8, c->cam.width, c->cam.height, c->stream, path);
}

int tpd_present_rgba8(tpd_ctx* c, void* d_dst, uint64_t dst_pitch, void* stream) {
    if (!c) return fail(TPD_ERR_INVALID, "tpd_present_rgba8: null context");
    hipSetDevice(c->device);
    if (!c->rendered) return fail(TPD_ERR_STATE, "tpd_present_rgba8: nothing rendered");
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    return copy_pitched(c->d_rgba8, c->cam.width, c->cam.height, d_dst, dst_pitch, s, "tpd_present_rgba8");
}

int tpd_batch_save_png(tpd_ctx* c, uint32_t view, const char* path) {
    if (!c || !path) return fail(TPD_ERR_INVALID, "tpd_batch_save_png: bad argument");
    hipSetDevice(c->device);
    const uint8_t* d = nullptr;
    uint32_t w = 0, h = 0;
    int rc = batch_view_rgba8(c, view, &d, &w, &h);
    return rc != TPD_OK ? rc : save_device_rgba8(d, w, h, nullptr, path);
}

int tpd_batch_present_rgba8(tpd_ctx* c, uint32_t view, void* d_dst, uint64_t dst_pitch, void* stream) {
    if (!c) return fail(TPD_ERR_INVALID, "tpd_batch_present_rgba8: null context");
    hipSetDevice(c->device);
    const uint8_t* d = nullptr;
    uint32_t w = 0, h = 0;
    int rc = batch_view_rgba8(c, view, &d, &w, &h);
    if (rc != TPD_OK) return rc;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    return copy_pitched(d, w, h, d_dst, dst_pitch, s, "tpd_batch_present_rgba8");
}

}  // extern "C"
