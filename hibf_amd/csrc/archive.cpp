// Archives of IBFs and HIBFs in the byte layout of a cereal BinaryArchive of
// the reference's objects (interleaved_bloom_filter.hpp:471-485,
// bit_vector.hpp:885-924, hierarchical_interleaved_bloom_filter.hpp:267-277,
// :300-309).
//
// The cereal binary-archive rules this layout follows: an arithmetic value is
// its native (little-endian) bytes; make_nvp(name, x) is x alone; base_class<B>
// is B's fields in place; binary_data(p, n) is n raw bytes; a std::vector of an
// arithmetic type is a uint64 element count followed by the elements; a
// std::vector of a class is the uint64 count followed by each element's
// fields.  None of these serialize functions takes a version argument, so the
// archive holds no class-version records.
//
// Bit matrices are streamed between the file and HBM in chunks, so an HIBF of
// tens of GB never needs a host copy of its arena.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "hibf_gpu.h"

namespace hibf_amd
{
int set_error(int status, std::string msg);
}

namespace
{

using hibf_amd::set_error;

constexpr uint32_t kArchiveVersion = 1;            // interleaved_bloom_filter.hpp:463, hierarchical_...hpp:258
constexpr uint64_t kChunkWords = uint64_t{1} << 23; // 64 MiB per transfer

struct File
{
    std::FILE * f{nullptr};
    ~File()
    {
        if (f)
            std::fclose(f);
    }
};

// ------------------------------------------------------------- writing --

struct Writer
{
    std::FILE * f;
    bool ok{true};

    void bytes(void const * p, size_t n)
    {
        if (ok && n && std::fwrite(p, 1, n, f) != n)
            ok = false;
    }
    void u64(uint64_t v) { bytes(&v, 8); }
    void u32(uint32_t v) { bytes(&v, 4); }
    void u8(uint8_t v) { bytes(&v, 1); }
    void u64s(uint64_t const * p, uint64_t n) { bytes(p, n * 8u); }
};

// Device words [d, d + n) to the file, through a host buffer.
int write_device_words(Writer & w, uint64_t const * d, uint64_t n, std::vector<uint64_t> & buf)
{
    for (uint64_t i = 0; i < n && w.ok; i += kChunkWords)
    {
        uint64_t const m = std::min(kChunkWords, n - i);
        buf.resize(m);
        hipError_t const e = hipMemcpy(buf.data(), d + i, m * 8u, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            return set_error(HIBF_GPU_ERR_DEVICE, std::string("archive download: ") + hipGetErrorString(e));
        }
        w.u64s(buf.data(), m);
    }
    return HIBF_GPU_OK;
}

// interleaved_bloom_filter::serialize for one IBF whose words are at d_words.
int write_ibf(Writer & w,
              hibf_gpu_ibf_geometry const & g,
              uint64_t const * d_words,
              uint64_t const * occupancy, // technical_bins entries, or NULL for zeros
              bool track_occupancy,
              std::vector<uint64_t> & buf)
{
    w.u32(kArchiveVersion);
    w.u64(g.bins);
    w.u64(g.technical_bins);
    w.u64(g.bin_size);
    w.u64(g.hash_shift);
    w.u64(g.bin_words);
    w.u64(g.hash_funs);
    w.u64(g.technical_bins * g.bin_size); // bit_vector size in bits
    if (int const rc = write_device_words(w, d_words, g.n_words, buf))
        return rc;
    w.u64(g.technical_bins);
    if (occupancy)
        w.u64s(occupancy, g.technical_bins);
    else
        for (uint64_t i = 0; i < g.technical_bins; ++i)
            w.u64(0u);
    w.u8(track_occupancy ? 1u : 0u);
    return HIBF_GPU_OK;
}

int open_for_write(File & file, char const * path, int append)
{
    if (!path)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null path");
    file.f = std::fopen(path, append ? "ab" : "wb");
    if (!file.f)
        return set_error(HIBF_GPU_ERR_IO, std::string("cannot open for writing: ") + path);
    return HIBF_GPU_OK;
}

int finish_write(File & file, Writer const & w, char const * path)
{
    bool ok = w.ok;
    if (std::fclose(file.f) != 0)
        ok = false;
    file.f = nullptr;
    return ok ? HIBF_GPU_OK : set_error(HIBF_GPU_ERR_IO, std::string("write failed: ") + path);
}

// ------------------------------------------------------------- reading --

struct Reader
{
    std::FILE * f;
    uint64_t pos;  // current byte offset
    uint64_t size; // file size
    bool ok{true};

    bool bytes(void * p, uint64_t n)
    {
        if (!ok || n > size - pos || (n && std::fread(p, 1, n, f) != n))
            return ok = false;
        pos += n;
        return true;
    }
    uint64_t u64()
    {
        uint64_t v = 0;
        bytes(&v, 8);
        return v;
    }
    uint32_t u32()
    {
        uint32_t v = 0;
        bytes(&v, 4);
        return v;
    }
    uint8_t u8()
    {
        uint8_t v = 0;
        bytes(&v, 1);
        return v;
    }
    bool skip(uint64_t n)
    {
        if (!ok || n > size - pos || fseeko(f, static_cast<off_t>(pos + n), SEEK_SET) != 0)
            return ok = false;
        pos += n;
        return true;
    }
    bool seek(uint64_t to)
    {
        if (to > size || fseeko(f, static_cast<off_t>(to), SEEK_SET) != 0)
            return ok = false;
        pos = to;
        return true;
    }
    // A vector length whose elements of `elem` bytes must still fit the file.
    bool length(uint64_t & n, uint64_t elem)
    {
        n = u64();
        if (ok && n > (size - pos) / elem)
            ok = false;
        return ok;
    }
};

int open_for_read(File & file, Reader & r, char const * path, uint64_t offset)
{
    if (!path)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null path");
    file.f = std::fopen(path, "rb");
    if (!file.f)
        return set_error(HIBF_GPU_ERR_IO, std::string("cannot open: ") + path);
    if (fseeko(file.f, 0, SEEK_END) != 0)
        return set_error(HIBF_GPU_ERR_IO, std::string("cannot seek: ") + path);
    off_t const end = ftello(file.f);
    r.f = file.f;
    r.size = end < 0 ? 0 : static_cast<uint64_t>(end);
    r.pos = 0;
    if (offset > r.size || !r.seek(offset))
        return set_error(HIBF_GPU_ERR_IO, "offset beyond the end of the file");
    return HIBF_GPU_OK;
}

int truncated(char const * what)
{
    return set_error(HIBF_GPU_ERR_IO, std::string("archive ends early or is malformed (") + what + ")");
}

// Header and placement of one IBF in a file; the words are read later.
struct IbfRecord
{
    hibf_gpu_ibf_geometry g{};
    uint64_t words_at{0}; // file offset of the bit_vector words
    std::vector<uint64_t> occupancy;
    bool track_occupancy{false};
};

uint64_t countl_zero64(uint64_t x)
{
    return x == 0 ? 64u : static_cast<uint64_t>(__builtin_clzll(x));
}

// Reads one interleaved_bloom_filter record.  The words are skipped (their
// file offset is kept for a later read); the occupancy is read when
// keep_occupancy is set and skipped otherwise.
int read_ibf_record(Reader & r, IbfRecord & rec, bool keep_occupancy)
{
    uint32_t const version = r.u32();
    hibf_gpu_ibf_geometry & g = rec.g;
    g.bins = r.u64();
    g.technical_bins = r.u64();
    g.bin_size = r.u64();
    g.hash_shift = r.u64();
    g.bin_words = r.u64();
    g.hash_funs = r.u64();
    uint64_t const bits = r.u64();
    if (!r.ok)
        return truncated("IBF header");
    if (version != kArchiveVersion)
        return set_error(HIBF_GPU_ERR_IO, "unsupported interleaved_bloom_filter archive version "
                                              + std::to_string(version));
    // The geometry relations of src/interleaved_bloom_filter.cpp:48-52.
    bool const consistent = g.bins > 0 && g.bin_size > 0 && g.hash_funs >= 1 && g.hash_funs <= 5
                         && g.bin_words == (g.bins + 63u) / 64u && g.technical_bins == g.bin_words * 64u
                         && g.hash_shift == countl_zero64(g.bin_size)
                         && g.bin_size <= UINT64_MAX / g.technical_bins && bits == g.technical_bins * g.bin_size;
    if (!consistent)
        return set_error(HIBF_GPU_ERR_IO, "archive fields are not a consistent interleaved_bloom_filter");
    g.n_words = bits / 64u;
    rec.words_at = r.pos;
    if (!r.skip(g.n_words * 8u))
        return truncated("bit matrix");
    uint64_t n_occ = 0;
    if (!r.length(n_occ, 8u))
        return truncated("occupancy");
    if (keep_occupancy)
    {
        if (n_occ != g.technical_bins)
            return set_error(HIBF_GPU_ERR_IO, "occupancy length differs from technical_bins");
        rec.occupancy.resize(n_occ);
        r.bytes(rec.occupancy.data(), n_occ * 8u);
    }
    else if (!r.skip(n_occ * 8u))
    {
        return truncated("occupancy");
    }
    rec.track_occupancy = r.u8() != 0;
    return r.ok ? HIBF_GPU_OK : truncated("track_occupancy");
}

// File words [at, at + n*8) to device words [d, d + n).
int read_words_to_device(Reader & r, uint64_t at, uint64_t * d, uint64_t n, std::vector<uint64_t> & buf)
{
    if (!r.seek(at))
        return truncated("bit matrix");
    for (uint64_t i = 0; i < n; i += kChunkWords)
    {
        uint64_t const m = std::min(kChunkWords, n - i);
        buf.resize(m);
        if (!r.bytes(buf.data(), m * 8u))
            return truncated("bit matrix");
        hipError_t const e = hipMemcpy(d + i, buf.data(), m * 8u, hipMemcpyHostToDevice);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            return set_error(HIBF_GPU_ERR_DEVICE, std::string("archive upload: ") + hipGetErrorString(e));
        }
    }
    return HIBF_GPU_OK;
}

// A vector<vector<uint64_t>> as flat values with offsets.
int read_nested(Reader & r, uint64_t expect_outer, std::vector<uint64_t> & flat, std::vector<uint64_t> & offs,
                char const * what)
{
    uint64_t n = 0;
    if (!r.length(n, 8u))
        return truncated(what);
    if (n != expect_outer)
        return set_error(HIBF_GPU_ERR_IO, std::string(what) + " has a different number of IBFs than ibf_vector");
    offs.assign(1, 0u);
    for (uint64_t i = 0; i < n; ++i)
    {
        uint64_t m = 0;
        if (!r.length(m, 8u))
            return truncated(what);
        size_t const at = flat.size();
        flat.resize(at + m);
        if (!r.bytes(flat.data() + at, m * 8u))
            return truncated(what);
        offs.push_back(flat.size());
    }
    return HIBF_GPU_OK;
}

} // namespace

extern "C" {

int hibf_gpu_ibf_save(hibf_gpu_ibf ibf, const char * path, int append)
{
    if (!ibf)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    hibf_gpu_ibf_geometry g{};
    uint64_t * d_words = nullptr;
    int track = 0;
    if (int rc = hibf_gpu_ibf_get_geometry(ibf, &g))
        return rc;
    if (int rc = hibf_gpu_ibf_device_words(ibf, &d_words))
        return rc;
    try
    {
        std::vector<uint64_t> occ(g.technical_bins);
        if (int rc = hibf_gpu_ibf_occupancy(ibf, occ.data(), &track))
            return rc;
        if (hipError_t e = hipDeviceSynchronize(); e != hipSuccess)
        {
            (void)hipGetLastError();
            return set_error(HIBF_GPU_ERR_DEVICE, std::string("archive sync: ") + hipGetErrorString(e));
        }
        File file;
        if (int rc = open_for_write(file, path, append))
            return rc;
        Writer w{file.f};
        std::vector<uint64_t> buf;
        if (int rc = write_ibf(w, g, d_words, occ.data(), track != 0, buf))
            return rc;
        return finish_write(file, w, path);
    }
    catch (...)
    {
        return set_error(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

int hibf_gpu_ibf_load(const char * path, uint64_t offset, hibf_gpu_ibf * out, uint64_t * end_offset)
{
    if (!out)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    try
    {
        File file;
        Reader r{};
        if (int rc = open_for_read(file, r, path, offset))
            return rc;
        IbfRecord rec;
        if (int rc = read_ibf_record(r, rec, true))
            return rc;
        uint64_t const end = r.pos;
        hibf_gpu_ibf ibf = nullptr;
        if (int rc = hibf_gpu_ibf_create(rec.g.bins, rec.g.bin_size, static_cast<uint32_t>(rec.g.hash_funs), &ibf))
            return rc;
        std::unique_ptr<std::remove_pointer_t<hibf_gpu_ibf>, int (*)(hibf_gpu_ibf)> guard{ibf, hibf_gpu_ibf_destroy};
        uint64_t * d_words = nullptr;
        if (int rc = hibf_gpu_ibf_device_words(ibf, &d_words))
            return rc;
        std::vector<uint64_t> buf;
        if (int rc = read_words_to_device(r, rec.words_at, d_words, rec.g.n_words, buf))
            return rc;
        if (int rc = hibf_gpu_ibf_set_occupancy(ibf, rec.occupancy.data(), rec.track_occupancy ? 1 : 0))
            return rc;
        if (end_offset)
            *end_offset = end;
        *out = guard.release();
        return HIBF_GPU_OK;
    }
    catch (...)
    {
        return set_error(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

int hibf_gpu_hibf_save(hibf_gpu_hibf hibf, const char * path, int append)
{
    if (!hibf)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    try
    {
        uint64_t n_ibfs = 0, n_tbs = 0, n_words = 0, n_ub = 0, depth = 0;
        if (int rc = hibf_gpu_hibf_info(hibf, &n_ibfs, &n_tbs, &n_words, &n_ub, &depth))
            return rc;
        std::vector<hibf_gpu_ibf_geometry> geoms(n_ibfs);
        std::vector<uint64_t> word_off(n_ibfs), tb_off(n_ibfs + 1), next(n_tbs), ub(n_tbs);
        if (int rc = hibf_gpu_hibf_export_tables(hibf, geoms.data(), word_off.data(), tb_off.data(), next.data(),
                                                 ub.data()))
            return rc;
        uint64_t * d_arena = nullptr;
        if (int rc = hibf_gpu_hibf_device_arena(hibf, &d_arena))
            return rc;
        // prev_ibf_id as src/hierarchical_interleaved_bloom_filter.cpp:130 fills it.
        std::vector<std::pair<uint64_t, uint64_t>> prev(n_ibfs, {0u, 0u});
        for (uint64_t i = 0; i < n_ibfs; ++i)
            for (uint64_t b = tb_off[i]; b < tb_off[i + 1]; ++b)
                if (ub[b] == HIBF_GPU_BIN_MERGED && next[b] < n_ibfs)
                    prev[next[b]] = {i, b - tb_off[i]};
        if (hipError_t e = hipDeviceSynchronize(); e != hipSuccess)
        {
            (void)hipGetLastError();
            return set_error(HIBF_GPU_ERR_DEVICE, std::string("archive sync: ") + hipGetErrorString(e));
        }
        File file;
        if (int rc = open_for_write(file, path, append))
            return rc;
        Writer w{file.f};
        std::vector<uint64_t> buf;
        w.u32(kArchiveVersion);
        w.u64(n_ub);
        w.u64(n_ibfs);
        for (uint64_t i = 0; i < n_ibfs; ++i)
            if (int rc = write_ibf(w, geoms[i], d_arena + word_off[i], nullptr, false, buf))
                return rc;
        w.u64(n_ibfs); // next_ibf_id
        for (uint64_t i = 0; i < n_ibfs; ++i)
        {
            w.u64(tb_off[i + 1] - tb_off[i]);
            w.u64s(next.data() + tb_off[i], tb_off[i + 1] - tb_off[i]);
        }
        w.u64(n_ibfs); // prev_ibf_id
        for (auto const & [ibf_idx, bin_idx] : prev)
        {
            w.u64(ibf_idx);
            w.u64(bin_idx);
        }
        w.u64(n_ibfs); // ibf_bin_to_user_bin_id
        for (uint64_t i = 0; i < n_ibfs; ++i)
        {
            w.u64(tb_off[i + 1] - tb_off[i]);
            w.u64s(ub.data() + tb_off[i], tb_off[i + 1] - tb_off[i]);
        }
        return finish_write(file, w, path);
    }
    catch (...)
    {
        return set_error(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

int hibf_gpu_hibf_load(const char * path, uint64_t offset, hibf_gpu_hibf * out, uint64_t * end_offset)
{
    if (!out)
        return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    try
    {
        File file;
        Reader r{};
        if (int rc = open_for_read(file, r, path, offset))
            return rc;
        uint32_t const version = r.u32();
        uint64_t const n_ub = r.u64();
        uint64_t n_ibfs = 0;
        // An IBF record is at least 4 + 8 * 8 + 8 + 1 bytes.
        if (!r.length(n_ibfs, 77u))
            return truncated("HIBF header");
        if (version != kArchiveVersion)
            return set_error(HIBF_GPU_ERR_IO, "unsupported hierarchical_interleaved_bloom_filter archive version "
                                                  + std::to_string(version));
        if (n_ibfs == 0)
            return set_error(HIBF_GPU_ERR_IO, "archive holds an HIBF without IBFs");
        std::vector<IbfRecord> recs(n_ibfs);
        for (IbfRecord & rec : recs)
            if (int rc = read_ibf_record(r, rec, false))
                return rc;
        std::vector<uint64_t> next, next_off, ub, ub_off;
        if (int rc = read_nested(r, n_ibfs, next, next_off, "next_ibf_id"))
            return rc;
        uint64_t n_prev = 0;
        if (!r.length(n_prev, 16u) || n_prev != n_ibfs || !r.skip(n_prev * 16u))
            return r.ok ? set_error(HIBF_GPU_ERR_IO, "prev_ibf_id has a different number of IBFs than ibf_vector")
                        : truncated("prev_ibf_id");
        if (int rc = read_nested(r, n_ibfs, ub, ub_off, "ibf_bin_to_user_bin_id"))
            return rc;
        uint64_t const end = r.pos;
        for (uint64_t i = 0; i < n_ibfs; ++i)
            if (next_off[i + 1] - next_off[i] != recs[i].g.bins || ub_off[i + 1] - ub_off[i] != recs[i].g.bins)
                return set_error(HIBF_GPU_ERR_IO, "routing tables of an IBF differ from its bin_count");
        // Arena placement: each IBF 256-byte aligned.
        std::vector<hibf_gpu_ibf_geometry> geoms(n_ibfs);
        std::vector<uint64_t> word_off(n_ibfs);
        uint64_t at = 0;
        for (uint64_t i = 0; i < n_ibfs; ++i)
        {
            geoms[i] = recs[i].g;
            word_off[i] = at;
            at += (recs[i].g.n_words + 31u) / 32u * 32u;
        }
        hibf_gpu_hibf h = nullptr;
        if (int rc = hibf_gpu_hibf_create(n_ibfs, geoms.data(), nullptr, word_off.data(), next_off.data(),
                                          next.data(), ub.data(), n_ub, &h))
            return rc;
        std::unique_ptr<std::remove_pointer_t<hibf_gpu_hibf>, int (*)(hibf_gpu_hibf)> guard{h, hibf_gpu_hibf_destroy};
        uint64_t * d_arena = nullptr;
        if (int rc = hibf_gpu_hibf_device_arena(h, &d_arena))
            return rc;
        std::vector<uint64_t> buf;
        for (uint64_t i = 0; i < n_ibfs; ++i)
            if (int rc = read_words_to_device(r, recs[i].words_at, d_arena + word_off[i], recs[i].g.n_words, buf))
                return rc;
        if (end_offset)
            *end_offset = end;
        *out = guard.release();
        return HIBF_GPU_OK;
    }
    catch (...)
    {
        return set_error(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

} // extern "C"
