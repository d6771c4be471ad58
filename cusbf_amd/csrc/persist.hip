// persist.hip -- filter images on disk (no reference counterpart: the reference
// keeps filters only in device memory).
//
// Layout (little-endian): a fixed 336-byte header -- magic "CUSBFIMG", format
// version, header size, the filter's cusbf_config, shard count, filter bits,
// word count and a checksum of the words -- then num_shards * 4 filter words
// in device order. The checksum (sum of a splitmix64 mix of each word and its
// index, modulo 2^64) is computed on the device before saving and after
// loading, so a bad image is found without a host pass over the words.
// Words move through a pinned bounce buffer in 64 MiB pieces.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/cusbf_hip.h"
#include "filter_internal.hpp"
#include "launch.hpp"

namespace {

namespace ci = cusbf_internal;

constexpr char kMagic[8] = {'C', 'U', 'S', 'B', 'F', 'I', 'M', 'G'};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kPiece = 64ull << 20;

struct ImageHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    cusbf_config config;
    uint32_t reserved;
    uint64_t num_shards;
    uint64_t filter_bits;
    uint64_t word_count;
    uint64_t checksum;
};
static_assert(sizeof(cusbf_config) == 284, "cusbf_config layout changed");
static_assert(sizeof(ImageHeader) == 336, "image header layout changed");

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

int io_error(const char* what, const char* path) {
    return ci::set_error(CUSBF_ERR_IO, (std::string(what) + path).c_str());
}

int device_checksum(const uint64_t* d_words, uint64_t n, uint64_t* out) {
    uint64_t* d_sum = nullptr;
    hipError_t e = hipMalloc(&d_sum, sizeof(uint64_t));
    if (e == hipSuccess) e = sbf::words_checksum(d_words, n, d_sum, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, d_sum, sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (d_sum) (void)hipFree(d_sum);
    return e == hipSuccess ? CUSBF_OK : ci::set_hip_error(e, __FILE__, __LINE__);
}

int read_header(FILE* fh, const char* path, ImageHeader& h) {
    if (std::fread(&h, sizeof(h), 1, fh) != 1 || std::memcmp(h.magic, kMagic, 8) != 0) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT,
                             (std::string("not a filter image: ") + path).c_str());
    }
    if (h.version != kVersion || h.header_bytes != sizeof(ImageHeader)) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT,
                             (std::string("unsupported filter image version: ") + path).c_str());
    }
    if (h.word_count != h.num_shards * 4 || h.filter_bits != h.num_shards * 256 || h.num_shards == 0 ||
        (h.num_shards & (h.num_shards - 1)) != 0) {
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT,
                             (std::string("corrupt filter image header: ") + path).c_str());
    }
    return CUSBF_OK;
}

}  // namespace

extern "C" {

int cusbf_save(const cusbf_filter* f, const char* path) {
    if (!f || !path) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t* d_words = nullptr;
    int rc = cusbf_words_device_ptr(f, &d_words);
    if (rc != CUSBF_OK) return rc;
    ImageHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(ImageHeader);
    h.config = ci::config(f);
    h.num_shards = cusbf_num_shards(f);
    h.filter_bits = cusbf_filter_bits(f);
    h.word_count = h.num_shards * 4;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
    rc = device_checksum(d_words, h.word_count, &h.checksum);
    if (rc != CUSBF_OK) return rc;

    File fh(std::fopen(path, "wb"));
    if (!fh) return io_error("Failed to open filter image for writing: ", path);
    if (std::fwrite(&h, sizeof(h), 1, fh.get()) != 1) return io_error("Failed to write filter image: ", path);
    char* bounce = nullptr;
    e = hipHostMalloc(&bounce, kPiece, hipHostMallocDefault);
    if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
    const uint64_t bytes = h.word_count * 8;
    for (uint64_t at = 0; at < bytes && rc == CUSBF_OK; at += kPiece) {
        const uint64_t n = bytes - at < kPiece ? bytes - at : kPiece;
        e = hipMemcpy(bounce, reinterpret_cast<const char*>(d_words) + at, n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = ci::set_hip_error(e, __FILE__, __LINE__);
        } else if (std::fwrite(bounce, 1, n, fh.get()) != n) {
            rc = io_error("Failed to write filter image: ", path);
        }
    }
    (void)hipHostFree(bounce);
    if (rc == CUSBF_OK && std::fflush(fh.get()) != 0) rc = io_error("Failed to write filter image: ", path);
    return rc;
}

int cusbf_image_config(const char* path, cusbf_config* config, uint64_t* filter_bits) {
    if (!path) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    File fh(std::fopen(path, "rb"));
    if (!fh) return io_error("Failed to open filter image: ", path);
    ImageHeader h{};
    const int rc = read_header(fh.get(), path, h);
    if (rc != CUSBF_OK) return rc;
    if (config) *config = h.config;
    if (filter_bits) *filter_bits = h.filter_bits;
    return CUSBF_OK;
}

int cusbf_load(const char* path, cusbf_filter** out) {
    if (!path || !out) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    File fh(std::fopen(path, "rb"));
    if (!fh) return io_error("Failed to open filter image: ", path);
    ImageHeader h{};
    int rc = read_header(fh.get(), path, h);
    if (rc != CUSBF_OK) return rc;
    cusbf_filter* f = nullptr;
    rc = cusbf_create(&h.config, h.filter_bits, &f);
    if (rc != CUSBF_OK) return rc;
    if (cusbf_num_shards(f) != h.num_shards) {
        cusbf_destroy(f);
        return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT,
                             (std::string("corrupt filter image header: ") + path).c_str());
    }
    uint64_t* d_words = nullptr;
    (void)cusbf_words_device_ptr(f, &d_words);
    char* bounce = nullptr;
    hipError_t e = hipHostMalloc(&bounce, kPiece, hipHostMallocDefault);
    if (e != hipSuccess) rc = ci::set_hip_error(e, __FILE__, __LINE__);
    const uint64_t bytes = h.word_count * 8;
    for (uint64_t at = 0; at < bytes && rc == CUSBF_OK; at += kPiece) {
        const uint64_t n = bytes - at < kPiece ? bytes - at : kPiece;
        if (std::fread(bounce, 1, n, fh.get()) != n) {
            rc = io_error("Failed to read filter image (truncated): ", path);
        } else {
            e = hipMemcpy(reinterpret_cast<char*>(d_words) + at, bounce, n, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = ci::set_hip_error(e, __FILE__, __LINE__);
        }
    }
    if (bounce) (void)hipHostFree(bounce);
    uint64_t sum = 0;
    if (rc == CUSBF_OK) rc = device_checksum(d_words, h.word_count, &sum);
    if (rc == CUSBF_OK && sum != h.checksum) {
        rc = io_error("filter image checksum mismatch: ", path);
    }
    if (rc != CUSBF_OK) {
        cusbf_destroy(f);
        return rc;
    }
    *out = f;
    return CUSBF_OK;
}

}  // extern "C"
