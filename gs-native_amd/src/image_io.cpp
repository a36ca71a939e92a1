// PNG / PPM codecs and the area downscale (see image_io.h).
#include "image_io.h"

#include "log.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <vector>

namespace gs_native {

namespace {

constexpr uint8_t kPngSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};

uint32_t BigEndian32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

void PutBigEndian32(std::vector<uint8_t>& out, uint32_t value) {
    for (int shift = 24; shift >= 0; shift -= 8) {
        out.push_back(static_cast<uint8_t>(value >> shift));
    }
}

bool ReadFile(const std::string& path, std::vector<uint8_t>& bytes) {
    std::ifstream file(path, std::ios::binary);
    if (!file) {
        return false;
    }
    bytes.assign(std::istreambuf_iterator<char>(file), std::istreambuf_iterator<char>());
    return true;
}

// PNG filter reconstruction of one scanline (filter types 0-4, bpp bytes per pixel).
bool Unfilter(uint8_t type, uint8_t* line, const uint8_t* previous, size_t length, int bpp) {
    auto paeth = [](int a, int b, int c) {
        int p = a + b - c;
        int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
        return pa <= pb && pa <= pc ? a : (pb <= pc ? b : c);
    };
    for (size_t i = 0; i < length; ++i) {
        int left = i >= static_cast<size_t>(bpp) ? line[i - bpp] : 0;
        int up = previous ? previous[i] : 0;
        int corner = previous && i >= static_cast<size_t>(bpp) ? previous[i - bpp] : 0;
        int add = 0;
        switch (type) {
        case 0: add = 0; break;
        case 1: add = left; break;
        case 2: add = up; break;
        case 3: add = (left + up) / 2; break;
        case 4: add = paeth(left, up, corner); break;
        default: return false;
        }
        line[i] = static_cast<uint8_t>(line[i] + add);
    }
    return true;
}

bool DecodePng(const std::vector<uint8_t>& bytes, const std::string& path, Image8& image) {
    size_t pos = 8;
    uint32_t width = 0, height = 0;
    int depth = 0, colour = -1, interlace = 0;
    std::vector<uint8_t> compressed;
    bool ended = false;
    while (pos + 12 <= bytes.size() && !ended) {
        uint32_t length = BigEndian32(&bytes[pos]);
        if (pos + 12 + static_cast<size_t>(length) > bytes.size()) {
            break;
        }
        const uint8_t* type = &bytes[pos + 4];
        const uint8_t* data = &bytes[pos + 8];
        if (std::memcmp(type, "IHDR", 4) == 0 && length >= 13) {
            width = BigEndian32(data);
            height = BigEndian32(data + 4);
            depth = data[8];
            colour = data[9];
            interlace = data[12];
        }
        else if (std::memcmp(type, "IDAT", 4) == 0) {
            compressed.insert(compressed.end(), data, data + length);
        }
        else if (std::memcmp(type, "IEND", 4) == 0) {
            ended = true;
        }
        pos += 12 + static_cast<size_t>(length);
    }
    int channels = colour == 0 ? 1 : colour == 2 ? 3 : colour == 4 ? 2 : colour == 6 ? 4 : 0;
    if (!ended || width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16) || depth != 8 ||
        channels == 0 || interlace != 0) {
        GS_NATIVE_ERROR("unsupported or damaged PNG %s (8-bit grey/RGB[A], non-interlaced only)", path.c_str());
        return false;
    }
    const size_t stride = static_cast<size_t>(width) * channels;
    std::vector<uint8_t> raw((stride + 1) * height);
    uLongf raw_size = static_cast<uLongf>(raw.size());
    if (uncompress(raw.data(), &raw_size, compressed.data(), static_cast<uLong>(compressed.size())) != Z_OK ||
        raw_size != raw.size()) {
        GS_NATIVE_ERROR("cannot inflate the image data of %s", path.c_str());
        return false;
    }
    image.width = static_cast<int>(width);
    image.height = static_cast<int>(height);
    image.data.assign(static_cast<size_t>(width) * height * 3, 0);
    for (uint32_t y = 0; y < height; ++y) {
        uint8_t* line = &raw[y * (stride + 1) + 1];
        const uint8_t* previous = y > 0 ? &raw[(y - 1) * (stride + 1) + 1] : nullptr;
        if (!Unfilter(raw[y * (stride + 1)], line, previous, stride, channels)) {
            GS_NATIVE_ERROR("invalid PNG filter in %s", path.c_str());
            return false;
        }
        for (uint32_t x = 0; x < width; ++x) {
            uint8_t* out = &image.data[(static_cast<size_t>(y) * width + x) * 3];
            const uint8_t* in = line + static_cast<size_t>(x) * channels;
            if (channels <= 2) {
                out[0] = out[1] = out[2] = in[0];
            }
            else {
                out[0] = in[0];
                out[1] = in[1];
                out[2] = in[2];
            }
        }
    }
    return true;
}

// Binary PPM: "P6" <whitespace / comments> width height maxval(255) <one whitespace> data.
bool DecodePpm(const std::vector<uint8_t>& bytes, const std::string& path, Image8& image) {
    size_t pos = 2;
    int values[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        while (pos < bytes.size() && (std::isspace(bytes[pos]) || bytes[pos] == '#')) {
            if (bytes[pos] == '#') {
                while (pos < bytes.size() && bytes[pos] != '\n') {
                    ++pos;
                }
            }
            else {
                ++pos;
            }
        }
        while (pos < bytes.size() && std::isdigit(bytes[pos]) && values[k] < (1 << 20)) {
            values[k] = values[k] * 10 + (bytes[pos++] - '0');
        }
    }
    ++pos;  // the single whitespace byte before the data
    const size_t size = static_cast<size_t>(values[0]) * values[1] * 3;
    if (values[0] <= 0 || values[1] <= 0 || values[2] != 255 || pos + size > bytes.size()) {
        GS_NATIVE_ERROR("unsupported or truncated PPM %s (P6, maxval 255 only)", path.c_str());
        return false;
    }
    image.width = values[0];
    image.height = values[1];
    image.data.assign(bytes.begin() + static_cast<std::ptrdiff_t>(pos),
        bytes.begin() + static_cast<std::ptrdiff_t>(pos + size));
    return true;
}

} // namespace

bool ReadImageRgb(const std::string& path, Image8& image) {
    std::vector<uint8_t> bytes;
    if (!ReadFile(path, bytes)) {
        GS_NATIVE_ERROR("cannot read image %s", path.c_str());
        return false;
    }
    if (bytes.size() >= 8 && std::memcmp(bytes.data(), kPngSignature, 8) == 0) {
        return DecodePng(bytes, path, image);
    }
    if (bytes.size() >= 2 && bytes[0] == 'P' && bytes[1] == '6') {
        return DecodePpm(bytes, path, image);
    }
    GS_NATIVE_ERROR("unsupported image format %s (PNG and binary PPM are read; JPEG needs a decoder this build does "
                    "not have)", path.c_str());
    return false;
}

bool WritePngRgb(const std::string& path, const Image8& image) {
    if (image.width <= 0 || image.height <= 0 ||
        image.data.size() != static_cast<size_t>(image.width) * image.height * 3) {
        return false;
    }
    const size_t stride = static_cast<size_t>(image.width) * 3;
    std::vector<uint8_t> raw((stride + 1) * image.height);
    for (int y = 0; y < image.height; ++y) {
        raw[y * (stride + 1)] = 0;  // filter type 0
        std::memcpy(&raw[y * (stride + 1) + 1], &image.data[y * stride], stride);
    }
    uLongf packed_size = compressBound(static_cast<uLong>(raw.size()));
    std::vector<uint8_t> packed(packed_size);
    if (compress2(packed.data(), &packed_size, raw.data(), static_cast<uLong>(raw.size()), 6) != Z_OK) {
        return false;
    }
    packed.resize(packed_size);
    std::vector<uint8_t> out(kPngSignature, kPngSignature + 8);
    auto chunk = [&out](const char* type, const uint8_t* data, size_t length) {
        PutBigEndian32(out, static_cast<uint32_t>(length));
        size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), data, data + length);
        PutBigEndian32(out, static_cast<uint32_t>(crc32(0L, &out[start], static_cast<uInt>(length + 4))));
    };
    uint8_t header[13];
    std::vector<uint8_t> fields;
    PutBigEndian32(fields, static_cast<uint32_t>(image.width));
    PutBigEndian32(fields, static_cast<uint32_t>(image.height));
    std::memcpy(header, fields.data(), 8);
    header[8] = 8;   // bit depth
    header[9] = 2;   // RGB
    header[10] = 0;  // deflate
    header[11] = 0;  // adaptive filtering
    header[12] = 0;  // no interlace
    chunk("IHDR", header, 13);
    chunk("IDAT", packed.data(), packed.size());
    chunk("IEND", nullptr, 0);
    std::ofstream file(path, std::ios::binary);
    file.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size()));
    return static_cast<bool>(file);
}

Image8 DownscaleArea(const Image8& image, int width, int height) {
    Image8 out;
    out.width = width;
    out.height = height;
    out.data.assign(static_cast<size_t>(width) * height * 3, 0);
    const double sx = static_cast<double>(image.width) / width;
    const double sy = static_cast<double>(image.height) / height;
    // Per output column: the input columns it covers and their weights (covered fractions).
    auto spans = [](int out_size, double scale, int in_size) {
        std::vector<std::vector<std::pair<int, double>>> result(static_cast<size_t>(out_size));
        for (int o = 0; o < out_size; ++o) {
            double begin = o * scale, end = std::min((o + 1) * scale, static_cast<double>(in_size));
            for (int i = static_cast<int>(std::floor(begin)); i < end && i < in_size; ++i) {
                double weight = std::min(end, i + 1.0) - std::max(begin, static_cast<double>(i));
                if (weight > 1e-12) {
                    result[static_cast<size_t>(o)].push_back({i, weight});
                }
            }
        }
        return result;
    };
    auto columns = spans(width, sx, image.width);
    auto rows = spans(height, sy, image.height);
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            double sum[3] = {0.0, 0.0, 0.0}, area = 0.0;
            for (const auto& [row, wy] : rows[static_cast<size_t>(y)]) {
                for (const auto& [column, wx] : columns[static_cast<size_t>(x)]) {
                    const uint8_t* in = &image.data[(static_cast<size_t>(row) * image.width + column) * 3];
                    for (int c = 0; c < 3; ++c) {
                        sum[c] += wx * wy * in[c];
                    }
                    area += wx * wy;
                }
            }
            for (int c = 0; c < 3; ++c) {
                double value = std::nearbyint(sum[c] / area);  // current rounding mode: half to even
                out.data[(static_cast<size_t>(y) * width + x) * 3 + c] =
                    static_cast<uint8_t>(std::min(255.0, std::max(0.0, value)));
            }
        }
    }
    return out;
}

} // namespace gs_native
