// fastx.hip -- FASTA/FASTQ ingest for the filter (host side of libcusbf_hip.so).
//
// Text -> records -> device, in chunks:
//  * byte sources: a caller buffer, an mmap'd file, a zlib stream (gzip files,
//    detected by their magic bytes like gzstreambuf.hpp:101-110) or a caller
//    read function;
//  * a line reader with std::getline semantics (lines end at '\n', one
//    trailing '\r' is dropped, text after the last '\n' is a line when it is
//    not empty) and a parser with the semantics and error texts of the
//    reference's streaming reader (Fastx.hpp:240-460): the first header fixes
//    FASTA ('>') or FASTQ ('@'); FASTA sequence lines run to the next '>'
//    line; FASTQ sequence lines run to a '+' line, then quality lines until
//    their total length reaches the sequence length; errors carry
//    "source:line:column";
//  * record sequences are appended back to back straight into pinned staging
//    (a record is padded to the symbol width with the alphabet separator), with
//    the record's first symbol in an offsets array -- the batched form's input,
//    so no separator normalisation pass runs;
//  * a chunk is flushed when the next record does not fit (or the record table
//    is full): H2D, one batched insert / query launch, the device-side report
//    counts per record (valid windows, positive windows), D2H of the counts and
//    hits. Two staging sets alternate, so the host parses the next chunk while
//    the device works on the previous one; per-record results are delivered in
//    source order once a chunk's event completes.
// Chunk size follows the reference's budget rule (filter.cuh:402-425,
// fastx_chunk.cuh:94-152): fill_fraction of free device memory minus a 64 MiB
// slack, divided over the chunk's device buffers, is the flush threshold (at
// most 64 MiB); a zero budget puts every record in a chunk of its own, and a
// record larger than the staging grows it (the reference processes such a
// record alone). Text held in memory is split into records by host threads
// (FASTA always, FASTQ where its record starts can be vouched for).
// A consumer returning non-zero stops the query (the filter stays usable).
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/cusbf_hip.h"
#include "filter_internal.hpp"
#include "host_pool.hpp"

namespace {

namespace ci = cusbf_internal;

// --- byte sources -----------------------------------------------------------------

class Source {
   public:
    virtual ~Source() = default;
    // Next bytes into dst (at most cap): the count, 0 at the end, -1 on a read error.
    virtual int64_t read(char* dst, uint64_t cap) = 0;
    // The whole input when it is already in memory (data() == nullptr otherwise).
    virtual std::string_view whole() const { return {}; }
};

class MemorySource final : public Source {
   public:
    MemorySource(const char* data, uint64_t bytes) : data_(data, bytes) {}
    int64_t read(char*, uint64_t) override { return 0; }
    std::string_view whole() const override {
        return data_.data() ? data_ : std::string_view("", 0);
    }

   private:
    std::string_view data_;
};

class MappedFile final : public Source {
   public:
    ~MappedFile() override {
        if (map_ && map_ != MAP_FAILED) munmap(map_, size_);
    }
    // 0 ok, else the failing step ("open", "stat", "mmap").
    const char* open(const char* path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return "open";
        struct stat st {};
        if (fstat(fd, &st) != 0) {
            close(fd);
            return "stat";
        }
        size_ = static_cast<uint64_t>(st.st_size);
        if (size_ > 0) {
            map_ = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
            if (map_ == MAP_FAILED) {
                close(fd);
                return "mmap";
            }
            madvise(map_, size_, MADV_SEQUENTIAL);
        }
        close(fd);
        return nullptr;
    }
    int64_t read(char*, uint64_t) override { return 0; }
    std::string_view whole() const override {
        return size_ ? std::string_view(static_cast<const char*>(map_), size_) : std::string_view("", 0);
    }

   private:
    void* map_ = nullptr;
    uint64_t size_ = 0;
};

class GzipSource final : public Source {
   public:
    ~GzipSource() override {
        if (gz_) gzclose(gz_);
    }
    bool open(const char* path) {
        gz_ = gzopen(path, "rb");
        if (gz_) gzbuffer(gz_, 1u << 20);
        return gz_ != nullptr;
    }
    int64_t read(char* dst, uint64_t cap) override {
        const unsigned n = static_cast<unsigned>(std::min<uint64_t>(cap, 1u << 30));
        const int got = gzread(gz_, dst, n);
        return got < 0 ? -1 : got;
    }

   private:
    gzFile gz_ = nullptr;
};

class CallbackSource final : public Source {
   public:
    CallbackSource(cusbf_read_fn fn, void* user) : fn_(fn), user_(user) {}
    int64_t read(char* dst, uint64_t cap) override { return fn_(user_, dst, cap); }

   private:
    cusbf_read_fn fn_;
    void* user_;
};

bool is_gzip(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    unsigned char magic[2] = {0, 0};
    const size_t n = std::fread(magic, 1, 2, f);
    std::fclose(f);
    return n == 2 && magic[0] == 0x1F && magic[1] == 0x8B;
}

// --- lines ----------------------------------------------------------------------------

// std::getline over a source: a line is the text up to '\n' (not included) with
// one trailing '\r' removed; the text after the last '\n' is a line when it is
// not empty. `unterminated()` tells whether the last line ended at the end of
// the input (the only case in which a failing getline leaves it in its buffer).
class LineReader {
   public:
    explicit LineReader(Source& src) : src_(src), mem_(src.whole()) {
        if (!mem_.data()) buf_.resize(1u << 22);
    }
    // 1: a line, 0: end of input, -1: read error.
    int next(std::string_view& line) {
        if (mem_.data()) {
            if (pos_ >= mem_.size()) return 0;
            const char* start = mem_.data() + pos_;
            const uint64_t left = mem_.size() - pos_;
            const void* nl = std::memchr(start, '\n', left);
            uint64_t len;
            if (nl) {
                len = static_cast<uint64_t>(static_cast<const char*>(nl) - start);
                pos_ += len + 1;
                unterminated_ = false;
            } else {
                len = left;
                pos_ = mem_.size();
                unterminated_ = true;
            }
            line = trim(std::string_view(start, len));
            return 1;
        }
        for (;;) {
            if (beg_ < end_) {
                const void* nl = std::memchr(buf_.data() + beg_, '\n', end_ - beg_);
                if (nl) {
                    const uint64_t len = static_cast<uint64_t>(static_cast<const char*>(nl) - (buf_.data() + beg_));
                    line = trim(std::string_view(buf_.data() + beg_, len));
                    beg_ += len + 1;
                    unterminated_ = false;
                    return 1;
                }
            }
            if (eof_) {
                if (beg_ >= end_) return 0;
                line = trim(std::string_view(buf_.data() + beg_, end_ - beg_));
                beg_ = end_;
                unterminated_ = true;
                return 1;
            }
            // compact, grow when one line fills the buffer, read more
            if (beg_ > 0) {
                std::memmove(buf_.data(), buf_.data() + beg_, end_ - beg_);
                end_ -= beg_;
                beg_ = 0;
            }
            if (end_ == buf_.size()) buf_.resize(buf_.size() * 2);
            const int64_t got = src_.read(buf_.data() + end_, buf_.size() - end_);
            if (got < 0) return -1;
            if (got == 0) eof_ = true;
            end_ += static_cast<uint64_t>(got);
        }
    }
    bool unterminated() const { return unterminated_; }

   private:
    static std::string_view trim(std::string_view s) {
        if (!s.empty() && s.back() == '\r') s.remove_suffix(1);
        return s;
    }
    Source& src_;
    std::string_view mem_;
    uint64_t pos_ = 0;
    std::vector<char> buf_;
    uint64_t beg_ = 0, end_ = 0;
    bool eof_ = false;
    bool unterminated_ = false;
};

// --- parser -----------------------------------------------------------------------------

// fastx_column_at (Fastx.hpp:261-268): 1-based column of byte_index, clamped to the line.
uint32_t column_at(std::string_view line, uint64_t byte_index) {
    if (line.empty()) return 1;
    return static_cast<uint32_t>(std::min<uint64_t>(byte_index, line.size() - 1) + 1);
}

struct ParseStatus {
    int code = 0;  // 0 ok, CUSBF_ERR_FASTX_PARSE, CUSBF_ERR_IO
    uint32_t line = 0, column = 0;
    const char* detail = nullptr;
};

// Receives a record's sequence piece by piece.
class SequenceSink {
   public:
    virtual ~SequenceSink() = default;
    virtual void append(std::string_view piece) = 0;
};

class Parser {
   public:
    explicit Parser(Source& src) : lines_(src) {}
    // A parser resuming inside a text: `lines_before` lines precede the source,
    // whose records are of `format` (1 FASTA, 2 FASTQ).
    Parser(Source& src, uint64_t lines_before, int format)
        : lines_(src), format_(format), line_no_(lines_before) {}

    // 1: a record (header into `header`, sequence into `sink`), 0: end, -1: error.
    int next(std::string& header, SequenceSink& sink, uint64_t& seq_len, ParseStatus& st) {
        std::string_view h;
        const int got = header_line(h, st);
        if (got <= 0) return got;
        const char tag = h.front();
        if (format_ == 0) {
            if (tag == '>') {
                format_ = 1;
            } else if (tag == '@') {
                format_ = 2;
            } else {
                return fail(st, "expected FASTA or FASTQ header", column_at(h, 0));
            }
        }
        if (tag != (format_ == 1 ? '>' : '@')) {
            return fail(st, "mixed FASTA and FASTQ records are not supported", column_at(h, 0));
        }
        header.assign(h.data() + 1, h.size() - 1);
        seq_len = 0;
        return format_ == 1 ? fasta_sequence(sink, seq_len, st) : fastq_sequence(sink, seq_len, st);
    }

   private:
    int fail(ParseStatus& st, const char* detail, uint32_t column) {
        st.code = CUSBF_ERR_FASTX_PARSE;
        st.line = static_cast<uint32_t>(line_no_);
        st.column = column;
        st.detail = detail;
        return -1;
    }
    int io_fail(ParseStatus& st) {
        st.code = CUSBF_ERR_IO;
        return -1;
    }
    // 1 line, 0 end, -1 io error; counts lines like the reference's getline loop
    int line(std::string_view& out, ParseStatus& st) {
        const int r = lines_.next(out);
        if (r < 0) return io_fail(st);
        if (r > 0) {
            ++line_no_;
            last_ = out;
        }
        return r;
    }
    int header_line(std::string_view& h, ParseStatus& st) {
        if (has_pending_) {
            has_pending_ = false;
            header_buf_.swap(pending_);
            h = header_buf_;
            return 1;
        }
        std::string_view l;
        for (;;) {
            const int r = line(l, st);
            if (r <= 0) return r;
            if (!l.empty()) {
                header_buf_.assign(l.data(), l.size());
                h = header_buf_;
                return 1;
            }
        }
    }
    int fasta_sequence(SequenceSink& sink, uint64_t& len, ParseStatus& st) {
        std::string_view l;
        for (;;) {
            const int r = line(l, st);
            if (r < 0) return -1;
            if (r == 0) return 1;
            if (!l.empty() && l.front() == '>') {
                pending_.assign(l.data(), l.size());
                has_pending_ = true;
                return 1;
            }
            sink.append(l);
            len += l.size();
        }
    }
    int fastq_sequence(SequenceSink& sink, uint64_t& len, ParseStatus& st) {
        std::string_view l;
        for (;;) {
            const int r = line(l, st);
            if (r < 0) return -1;
            if (r == 0) {
                // what getline leaves in its buffer: the last line only when the
                // input ended without a newline
                const std::string_view buf = lines_.unterminated() ? last_ : std::string_view();
                return fail(st, "unterminated FASTQ record: missing '+' separator",
                            column_at(buf, buf.size() > 0 ? buf.size() - 1 : 0));
            }
            if (!l.empty() && l.front() == '+') return fastq_qualities(len, st);
            sink.append(l);
            len += l.size();
        }
    }
    // Fastx.hpp:436-460 with the columns of :270-287
    int fastq_qualities(uint64_t expected, ParseStatus& st) {
        uint64_t got = 0;
        uint32_t short_column = 1;
        std::string_view l;
        while (got < expected) {
            const int r = line(l, st);
            if (r < 0) return -1;
            if (r == 0) break;
            short_column = static_cast<uint32_t>(l.empty() ? 1 : l.size() + 1);
            got += l.size();
            if (got > expected) {
                const uint64_t before = got - l.size();
                const uint64_t idx = expected > before ? expected - before : l.size();
                return fail(st, "FASTQ quality length exceeds sequence length", column_at(l, idx));
            }
        }
        if (got == expected) return 1;
        return fail(st, "FASTQ quality length does not match sequence length", short_column);
    }

    LineReader lines_;
    int format_ = 0;  // 0 unknown, 1 FASTA, 2 FASTQ
    uint64_t line_no_ = 0;
    std::string pending_;
    bool has_pending_ = false;
    std::string header_buf_;
    std::string_view last_;
};

// --- FASTA text in memory, split by host threads ---------------------------------------
//
// After its first header a FASTA text cannot fail to parse (Fastx.hpp:240-460:
// a line starting with '>' starts a record, every other line is sequence), so a
// text held in memory is split into records by the host pool: pass 1 finds the
// header lines and counts each record's sequence bytes in parts of the text,
// pass 2 copies the sequence lines to their offsets. Lines, '\r' trimming and
// headers are those of the line reader and parser above.

struct SplitRecord {
    uint64_t header_at, header_len;  // in the text: after '>', trailing '\r' dropped
    uint64_t bytes;                  // sequence bytes (before padding)
    uint64_t out_at;                 // offset in the output
};

// One chunk's worth of records: text [pos, cut), output bytes after padding.
struct SplitBatch {
    std::vector<SplitRecord> records;
    uint64_t cut = 0;
    uint64_t out_bytes = 0;
};

// Start of the first line that is not empty once one trailing '\r' is dropped
// (the parser's first header line), or text.size() when there is none.
uint64_t first_content_line(std::string_view t) {
    uint64_t at = 0;
    while (at < t.size()) {
        const void* nl = std::memchr(t.data() + at, '\n', t.size() - at);
        const uint64_t stop = nl ? static_cast<uint64_t>(static_cast<const char*>(nl) - t.data()) : t.size();
        uint64_t len = stop - at;
        if (len != 0 && t[at + len - 1] == '\r') --len;
        if (len != 0) return at;
        at = nl ? stop + 1 : t.size();
    }
    return t.size();
}

class FastaSplitter {
   public:
    FastaSplitter(std::string_view text, uint32_t width) : t_(text.data()), end_(text.size()), width_(width) {}

    // Records from the header line at `pos` (a line start whose line begins with
    // '>') holding at most about `limit` bytes of text and at most `max_records`
    // records; a record larger than `limit` is taken alone, whole.
    void scan(uint64_t pos, uint64_t limit, uint64_t max_records, SplitBatch* b) {
        uint64_t span_end = end_ - pos > limit ? pos + limit : end_;
        if (span_end < end_) span_end = line_start_at_or_after(std::max(span_end, pos + 1));
        parts_.clear();
        split(pos, span_end);
        pass1();
        std::vector<SplitRecord>& recs = b->records;
        recs.clear();
        collect(recs);
        uint64_t cut = span_end;
        if (span_end < end_) {
            if (recs.size() >= 2) {
                // the last record may continue past the span: it starts the next batch
                cut = recs.back().header_at - 1;
                recs.pop_back();
                clip_parts(cut);
            } else {
                // one record larger than the span: extend it to the next header
                const uint64_t next = next_header(span_end);
                const size_t first_new = parts_.size();
                split(span_end, next);
                pass1(first_new);
                cut = next;
            }
        }
        if (recs.size() > max_records) {
            cut = recs[max_records].header_at - 1;
            recs.resize(max_records);
            clip_parts(cut);
        }
        // totals: each part's leading bytes belong to the record before it
        recs.clear();
        collect(recs);
        uint64_t at = 0;
        for (SplitRecord& r : recs) {
            r.out_at = at;
            at += padded(r.bytes);
        }
        b->cut = cut;
        b->out_bytes = at;
    }

    // Pass 2: the batch's sequence bytes into `out` (padding is the caller's).
    void copy(const SplitBatch& b, char* out) {
        const SplitRecord* recs = b.records.data();
        uint64_t next = 0;  // global index of the next record to start
        uint64_t tail = 0;  // output position after the bytes so far of record next - 1
        for (Part& p : parts_) {
            p.lead_out = tail;  // zero lead bytes in the first part: the text starts at a header
            tail += p.lead;
            p.first_own = next;
            for (const SplitRecord& r : p.recs) {
                tail = recs[next].out_at + r.bytes;
                ++next;
            }
        }
        cusbf_internal::host_pool().run(parts_.size(), [&](size_t i) { copy_part(parts_[i], recs, out); });
    }

   private:
    struct Part {
        uint64_t a = 0, b = 0;  // line starts
        uint64_t lead = 0;      // sequence bytes before the part's first header
        std::vector<SplitRecord> recs;  // records whose header is in the part, bytes within it
        uint64_t first_own = 0;  // global index of recs[0]
        uint64_t lead_out = 0;   // output offset of the lead bytes
    };

    uint64_t padded(uint64_t bytes) const {
        const uint64_t rem = bytes % width_;
        return rem ? bytes + (width_ - rem) : bytes;
    }
    uint64_t line_start_at_or_after(uint64_t at) const {
        if (at == 0 || t_[at - 1] == '\n') return at;
        const void* nl = std::memchr(t_ + at, '\n', end_ - at);
        return nl ? static_cast<uint64_t>(static_cast<const char*>(nl) - t_) + 1 : end_;
    }
    // The first line start at or after `at` (itself a line start) whose line
    // begins with '>' (sequence lines rarely hold a '>', so memchr leaps to it).
    uint64_t next_header(uint64_t at) const {
        for (uint64_t from = at; from < end_;) {
            const void* hit = std::memchr(t_ + from, '>', end_ - from);
            if (!hit) break;
            const uint64_t h = static_cast<uint64_t>(static_cast<const char*>(hit) - t_);
            if (h == at || t_[h - 1] == '\n') return h;
            from = h + 1;
        }
        return end_;
    }
    // Parts of about 1 MiB and more, one per pool task, cut at line starts.
    void split(uint64_t a, uint64_t b) {
        if (a >= b) return;
        constexpr uint64_t kMinPart = 1ull << 20;
        const uint64_t n = std::clamp<uint64_t>((b - a) / kMinPart, 1, 64);
        uint64_t start = a;
        for (uint64_t i = 1; i <= n && start < b; ++i) {
            uint64_t stop = i == n ? b : std::min(b, line_start_at_or_after(a + (b - a) / n * i));
            if (stop <= start) continue;
            Part p;
            p.a = start;
            p.b = stop;
            parts_.push_back(std::move(p));
            start = stop;
        }
    }
    // A line [at, next) of the text: its length after dropping one trailing '\r'.
    uint64_t line_at(uint64_t at, uint64_t limit, uint64_t* len) const {
        const void* nl = std::memchr(t_ + at, '\n', limit - at);
        const uint64_t stop = nl ? static_cast<uint64_t>(static_cast<const char*>(nl) - t_) : limit;
        uint64_t l = stop - at;
        if (l != 0 && t_[at + l - 1] == '\r') --l;
        *len = l;
        return nl ? stop + 1 : limit;
    }
    void scan_part(Part& p) const {
        p.lead = 0;
        p.recs.clear();
        for (uint64_t at = p.a; at < p.b;) {
            uint64_t len = 0;
            const uint64_t next = line_at(at, p.b, &len);
            if (len != 0 && t_[at] == '>') {
                p.recs.push_back(SplitRecord{at + 1, len - 1, 0, 0});
            } else if (p.recs.empty()) {
                p.lead += len;
            } else {
                p.recs.back().bytes += len;
            }
            at = next;
        }
    }
    void pass1(size_t from = 0) {
        const size_t n = parts_.size() - from;
        cusbf_internal::host_pool().run(n, [&](size_t i) { scan_part(parts_[from + i]); });
    }
    // Records in text order with their whole sequence bytes.
    void collect(std::vector<SplitRecord>& recs) const {
        for (const Part& p : parts_) {
            if (!recs.empty()) recs.back().bytes += p.lead;
            for (const SplitRecord& r : p.recs) recs.push_back(r);
        }
    }
    // Drops the parts and records from `cut` (a header line start) on.
    void clip_parts(uint64_t cut) {
        while (!parts_.empty() && parts_.back().a >= cut) parts_.pop_back();
        if (parts_.empty()) return;
        Part& p = parts_.back();
        if (p.b > cut) {
            p.b = cut;
            while (!p.recs.empty() && p.recs.back().header_at - 1 >= cut) p.recs.pop_back();
        }
    }
    void copy_part(const Part& p, const SplitRecord* recs, char* out) const {
        uint64_t rec = p.first_own;
        uint64_t to = p.lead_out;
        bool started = false;
        for (uint64_t at = p.a; at < p.b;) {
            uint64_t len = 0;
            const uint64_t next = line_at(at, p.b, &len);
            if (len != 0 && t_[at] == '>') {
                if (started) ++rec;
                started = true;
                to = recs[rec].out_at;
            } else if (len != 0) {
                std::memcpy(out + to, t_ + at, len);
                to += len;
            }
            at = next;
        }
    }

    const char* t_;
    uint64_t end_;
    uint32_t width_;
    std::vector<Part> parts_;
};

// --- FASTQ text in memory, split by host threads ---------------------------------------
//
// A FASTQ text can fail anywhere, so its parts are parsed by the serial rules
// above (blank lines before a header skipped; sequence lines up to a '+' line;
// quality lines until their bytes reach the sequence's) and checked: the parts
// start at record starts found by a pattern ('@' line, one sequence line, '+'
// line, a quality line as long as the sequence, then an '@' line or the end),
// and every part must end exactly where the next one starts. The first part
// starts where the serial parser is between records, and a part that ends on
// the next part's start leaves it there too, so every part then parses as the
// serial parser would. Anything else -- an error, a record running past its
// part, no record start near a chunk's end -- hands the text from the chunk's
// start to the serial parser, which reports errors at their line and column.

class FastqSplitter {
   public:
    FastqSplitter(std::string_view text, uint32_t width) : t_(text.data()), end_(text.size()), width_(width) {}

    // Records from the record start `pos`, about `limit` bytes of text and at most
    // `max_records` of them; false when the text from `pos` is left to the serial parser.
    bool scan(uint64_t pos, uint64_t limit, uint64_t max_records, SplitBatch* b) {
        uint64_t span_end = end_;
        if (end_ - pos > limit) {
            span_end = record_start_near(std::max(pos + limit, pos + 1));
            if (span_end == kNone) return false;
        }
        parts_.clear();
        constexpr uint64_t kMinPart = 1ull << 20;
        const uint64_t n = std::clamp<uint64_t>((span_end - pos) / kMinPart, 1, 64);
        uint64_t start = pos;
        for (uint64_t i = 1; i <= n; ++i) {
            uint64_t stop = span_end;
            if (i < n) {
                stop = record_start_near(pos + (span_end - pos) / n * i);
                if (stop == kNone || stop <= start || stop >= span_end) continue;
            }
            Part p;
            p.a = start;
            p.b = stop;
            parts_.push_back(std::move(p));
            start = stop;
        }
        cusbf_internal::host_pool().run(parts_.size(), [&](size_t i) { parse_part(parts_[i]); });
        std::vector<SplitRecord>& recs = b->records;
        recs.clear();
        for (const Part& p : parts_) {
            if (!p.ok) return false;
            recs.insert(recs.end(), p.recs.begin(), p.recs.end());
        }
        uint64_t cut = span_end;
        if (recs.size() > max_records) {
            cut = recs[max_records].header_at - 1;
            recs.resize(max_records);
            while (!parts_.empty() && parts_.back().a >= cut) parts_.pop_back();
            if (parts_.empty()) return false;  // cannot happen: cut lies past the first record
            Part& last = parts_.back();
            last.b = std::min(last.b, cut);
            while (!last.recs.empty() && last.recs.back().header_at - 1 >= cut) last.recs.pop_back();
        }
        uint64_t at = 0, next = 0;
        for (Part& p : parts_) {
            p.first_own = next;
            next += p.recs.size();
        }
        for (SplitRecord& r : recs) {
            r.out_at = at;
            const uint64_t rem = r.bytes % width_;
            at += rem ? r.bytes + (width_ - rem) : r.bytes;
        }
        b->cut = cut;
        b->out_bytes = at;
        return true;
    }

    // The batch's sequence bytes into `out` (padding is the caller's).
    void copy(const SplitBatch& b, char* out) {
        cusbf_internal::host_pool().run(parts_.size(), [&](size_t i) { copy_part(parts_[i], b.records.data(), out); });
    }

   private:
    static constexpr uint64_t kNone = ~0ull;
    static constexpr uint64_t kSearch = 4ull << 20;  // how far a record start is looked for

    struct Part {
        uint64_t a = 0, b = 0;
        std::vector<SplitRecord> recs;
        uint64_t first_own = 0;
        bool ok = false;
    };

    uint64_t line_at(uint64_t at, uint64_t* len) const {
        const void* nl = std::memchr(t_ + at, '\n', end_ - at);
        const uint64_t stop = nl ? static_cast<uint64_t>(static_cast<const char*>(nl) - t_) : end_;
        uint64_t l = stop - at;
        if (l != 0 && t_[at + l - 1] == '\r') --l;
        *len = l;
        return nl ? stop + 1 : end_;
    }

    // A line start at or after `x` where a record by the pattern begins; end_ when
    // the text ends first, kNone when none is found within kSearch bytes.
    uint64_t record_start_near(uint64_t x) const {
        uint64_t at = x;
        if (at < end_ && at > 0 && t_[at - 1] != '\n') {
            const void* nl = std::memchr(t_ + at, '\n', end_ - at);
            at = nl ? static_cast<uint64_t>(static_cast<const char*>(nl) - t_) + 1 : end_;
        }
        while (at < end_) {
            if (at - x > kSearch) return kNone;
            uint64_t l0 = 0;
            const uint64_t a1 = line_at(at, &l0);
            if (l0 != 0 && t_[at] == '@' && a1 < end_) {
                uint64_t l1 = 0, l2 = 0, l3 = 0;
                const uint64_t a2 = line_at(a1, &l1);
                if (a2 < end_ && !(l1 != 0 && t_[a1] == '+')) {
                    const uint64_t a3 = line_at(a2, &l2);
                    if (l2 != 0 && t_[a2] == '+' && a3 < end_) {
                        uint64_t a4 = line_at(a3, &l3);
                        if (l3 == l1) {
                            uint64_t l4 = 0;
                            while (a4 < end_) {  // blank lines, then a header or the end
                                const uint64_t a5 = line_at(a4, &l4);
                                if (l4 != 0) break;
                                a4 = a5;
                            }
                            if (a4 >= end_ || t_[a4] == '@') return at;
                        }
                    }
                }
            }
            at = a1;
        }
        return end_;
    }

    // Pass 1: the part's records by the serial rules; ok when it ends on p.b.
    void parse_part(Part& p) const {
        p.recs.clear();
        p.ok = false;
        uint64_t at = p.a;
        while (at < p.b) {
            uint64_t len = 0;
            uint64_t next = line_at(at, &len);
            if (len == 0) {
                at = next;
                continue;
            }
            if (t_[at] != '@') return;  // an error of the serial parser
            SplitRecord r{at + 1, len - 1, 0, 0};
            at = next;
            for (;;) {  // sequence lines up to the '+' line
                if (at >= p.b) return;
                next = line_at(at, &len);
                const bool plus = len != 0 && t_[at] == '+';
                at = next;
                if (plus) break;
                r.bytes += len;
            }
            uint64_t got = 0;
            while (got < r.bytes) {  // quality lines
                if (at >= p.b) return;
                at = line_at(at, &len);
                got += len;
            }
            if (got != r.bytes) return;
            p.recs.push_back(r);
        }
        p.ok = at == p.b;
    }

    // Pass 2: each record's sequence lines (from the line after its header up to
    // its '+' line) to the record's offset.
    void copy_part(const Part& p, const SplitRecord* recs, char* out) const {
        for (size_t i = 0; i < p.recs.size(); ++i) {
            const SplitRecord& own = p.recs[i];
            char* to = out + recs[p.first_own + i].out_at;
            uint64_t len = 0;
            uint64_t at = line_at(own.header_at + own.header_len, &len);  // past the header line
            for (;;) {
                const uint64_t next = line_at(at, &len);
                if (len != 0 && t_[at] == '+') break;
                std::memcpy(to, t_ + at, len);
                to += len;
                at = next;
            }
        }
    }

    const char* t_;
    uint64_t end_;
    uint32_t width_;
    std::vector<Part> parts_;
};

// Lines before `bytes` of text (each ends with '\n'), counted on the host pool.
uint64_t count_newlines(const char* t, uint64_t bytes) {
    constexpr uint64_t kPart = 16ull << 20;
    const uint64_t parts = (bytes + kPart - 1) / kPart;
    std::vector<uint64_t> counts(parts, 0);
    ci::host_pool().run(parts, [&](size_t i) {
        const char* a = t + i * kPart;
        const char* e = t + std::min(bytes, (i + 1) * kPart);
        uint64_t n = 0;
        while (a < e) {
            const void* nl = std::memchr(a, '\n', static_cast<size_t>(e - a));
            if (!nl) break;
            ++n;
            a = static_cast<const char*>(nl) + 1;
        }
        counts[i] = n;
    });
    uint64_t total = 0;
    for (uint64_t c : counts) total += c;
    return total;
}


// --- chunk pipeline -------------------------------------------------------------------

struct RecordMeta {
    uint64_t index;
    uint64_t header_at, header_len;
    uint64_t byte_at, bytes;  // in the chunk's staging
};

struct Chunk {
    char* h_seq = nullptr;
    uint64_t used = 0;  // bytes
    uint64_t* h_off = nullptr;     // records + 1 symbol offsets
    uint64_t* h_req = nullptr;     // per record: first window, window count
    uint64_t* h_counts = nullptr;  // per record: valid windows, positive windows
    uint8_t* h_hits = nullptr;
    hipEvent_t done = nullptr;
    bool in_flight = false;
    std::vector<RecordMeta> records;
    std::string headers;
};

constexpr uint64_t kSlackBytes = 64ull << 20;
constexpr uint64_t kDefaultMaxChunk = 64ull << 20;
constexpr uint64_t kMinStaging = 1ull << 20;
constexpr uint64_t kBytesPerRecordSlot = 32;  // staging bytes per record-table entry

class Pipeline final : public SequenceSink {
   public:
    Pipeline(cusbf_filter* f, bool query, hipStream_t stream, cusbf_fastx_consumer consumer,
             void* user, std::string source)
        : f_(f), query_(query), stream_(stream), consumer_(consumer), user_(user),
          source_(std::move(source)), width_(ci::symbol_width(f)), k_(ci::kmer_length(f)),
          sep_(static_cast<char>(ci::separator(f))) {}

    ~Pipeline() override {
        (void)hipStreamSynchronize(stream_);
        free_buffers();
        for (Chunk& c : chunks_) {
            if (c.done) (void)hipEventDestroy(c.done);
        }
    }

    // Flush threshold from fill_fraction (or the filter's chunk override), then
    // both staging sets and the device buffers.
    int init(double fill_fraction) {
        uint64_t target = ci::fastx_chunk_bytes(f_);
        if (target == 0) {
            size_t free_b = 0, total_b = 0;
            const hipError_t e = hipMemGetInfo(&free_b, &total_b);
            if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
            const double avail = free_b > kSlackBytes ? static_cast<double>(free_b - kSlackBytes) : 0.0;
            const uint64_t budget = fill_fraction > 0.0 ? static_cast<uint64_t>(avail * fill_fraction) : 0;
            // per chunk byte: sequence + start bitmap + record table (+ hits)
            const double per_byte = 1.0 + 0.125 + 40.0 / kBytesPerRecordSlot + (query_ ? 1.0 : 0.0);
            target = std::min(static_cast<uint64_t>(static_cast<double>(budget) / per_byte), kDefaultMaxChunk);
        }
        target_ = target;
        for (Chunk& c : chunks_) {
            const hipError_t e = hipEventCreateWithFlags(&c.done, hipEventDisableTiming);
            if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
        }
        return allocate(std::max(target_, kMinStaging));
    }

    int run(Source& src) {
        const std::string_view text = src.whole();
        if (text.data() != nullptr) {
            const uint64_t first = first_content_line(text);
            if (first < text.size() && text[first] == '>') return run_fasta_text(text, first);
            if (first < text.size() && text[first] == '@') {
                uint64_t resume = 0, index = 0;
                const int rc = run_fastq_text(text, first, &resume, &index);
                if (rc != CUSBF_OK || stopped_ || resume >= text.size()) return finish_run(rc);
                // the rest by the serial parser, with the lines before it counted
                MemorySource rest(text.data() + resume, text.size() - resume);
                const uint64_t lines = count_newlines(text.data(), resume);
                Parser parser(rest, lines, 2);
                return run_serial(parser, index);
            }
        }
        Parser parser(src);
        return run_serial(parser, 0);
    }

    // FASTQ text in memory: chunks split by the host pool while the splitter can
    // vouch for them; *resume = where the serial parser takes over (text.size()
    // when nothing is left), *index = the next record's index.
    int run_fastq_text(std::string_view text, uint64_t pos, uint64_t* resume, uint64_t* index) {
        FastqSplitter sp(text, width_);
        SplitBatch b;
        while (pos < text.size() && !stopped_) {
            if (!sp.scan(pos, target_, max_records_, &b)) break;
            int rc = CUSBF_OK;
            if (b.out_bytes > alloc_) {
                rc = finish(chunks_[cur_ ^ 1]);
                if (rc == CUSBF_OK) rc = allocate(b.out_bytes);
                if (rc != CUSBF_OK) return rc;
            }
            Chunk& c = chunks_[cur_];
            sp.copy(b, c.h_seq);
            rc = add_split_records(c, text, b, index);
            if (rc == CUSBF_OK) rc = flush();
            if (rc == CUSBF_OK) rc = take_next();
            if (rc != CUSBF_OK) return rc;
            pos = b.cut;
        }
        *resume = pos;
        return CUSBF_OK;
    }

    // A split batch's records into chunk c (its staging already holds their bytes):
    // separator padding, record table entries, headers for a consumer, totals.
    int add_split_records(Chunk& c, std::string_view text, const SplitBatch& b, uint64_t* index) {
        c.used = b.out_bytes;
        for (const SplitRecord& r : b.records) {
            const uint64_t pad = padded(r.bytes) - r.bytes;
            if (pad) std::memset(c.h_seq + r.out_at + r.bytes, sep_, pad);
            RecordMeta m{(*index)++, 0, 0, r.out_at, r.bytes};
            if (query_ && consumer_) {
                m.header_at = c.headers.size();
                m.header_len = r.header_len;
                c.headers.append(text.data() + r.header_at, r.header_len);
            }
            c.records.push_back(m);
            ++records_;
            bases_ += r.bytes;
        }
        return CUSBF_OK;
    }

    // After the last chunk: the device work drained, the consumer's stop reported.
    int finish_run(int rc) {
        if (rc != CUSBF_OK) return rc;
        rc = stopped_ ? CUSBF_OK : flush();
        if (rc == CUSBF_OK) rc = drain();
        if (rc == CUSBF_OK && stopped_) {
            return ci::set_error(CUSBF_ERR_IO, "FASTX record consumer stopped the query");
        }
        return rc;
    }

    int run_serial(Parser& parser, uint64_t first_index) {
        std::string header;
        ParseStatus st;
        for (uint64_t index = first_index;; ++index) {
            Chunk& c = chunks_[cur_];
            rec_at_ = c.used;
            spilled_ = false;
            spill_.clear();
            uint64_t len = 0;
            const int got = parser.next(header, *this, len, st);
            if (got < 0) {
                const int rc = drain();
                if (rc != CUSBF_OK) return rc;
                if (st.code == CUSBF_ERR_IO) {
                    return ci::set_error(CUSBF_ERR_IO,
                                         ("Failed to read FASTA/FASTQ input from " + source_).c_str());
                }
                return ci::set_fastx_error(source_.c_str(), st.line, st.column, st.detail);
            }
            if (got == 0) break;
            const int rc = close_record(index, header, len);
            if (rc != CUSBF_OK) return rc;
            if (stopped_) break;
        }
        return finish_run(CUSBF_OK);
    }

    // FASTA text in memory: each chunk's records split out by the host pool
    // straight into the chunk's pinned staging.
    int run_fasta_text(std::string_view text, uint64_t pos) {
        FastaSplitter sp(text, width_);
        SplitBatch b;
        uint64_t index = 0;
        while (pos < text.size() && !stopped_) {
            sp.scan(pos, target_, max_records_, &b);
            int rc = CUSBF_OK;
            if (b.out_bytes > alloc_) {  // one record larger than the staging
                rc = finish(chunks_[cur_ ^ 1]);
                if (rc == CUSBF_OK) rc = allocate(b.out_bytes);
                if (rc != CUSBF_OK) return rc;
            }
            Chunk& c = chunks_[cur_];
            sp.copy(b, c.h_seq);
            rc = add_split_records(c, text, b, &index);
            if (rc == CUSBF_OK) rc = flush();
            if (rc == CUSBF_OK) rc = take_next();
            if (rc != CUSBF_OK) return rc;
            pos = b.cut;
        }
        return finish_run(CUSBF_OK);
    }

    // Sequence bytes of the record being parsed go to the current chunk while
    // they fit, otherwise to the spill string (moved to the next chunk when the
    // record ends).
    void append(std::string_view piece) override {
        Chunk& c = chunks_[cur_];
        if (!spilled_ && c.used + piece.size() <= alloc_) {
            std::memcpy(c.h_seq + c.used, piece.data(), piece.size());
            c.used += piece.size();
            return;
        }
        if (!spilled_) {
            spill_.assign(c.h_seq + rec_at_, c.used - rec_at_);
            c.used = rec_at_;
            spilled_ = true;
        }
        spill_.append(piece.data(), piece.size());
    }

    cusbf_insert_report insert_report() const {
        return cusbf_insert_report{records_, bases_, valid_};
    }
    cusbf_query_report query_report() const {
        return cusbf_query_report{records_, bases_, valid_, positive_};
    }

   private:
    uint64_t padded(uint64_t bytes) const {
        const uint64_t rem = bytes % width_;
        return rem ? bytes + (width_ - rem) : bytes;
    }

    int close_record(uint64_t index, const std::string& header, uint64_t len) {
        const uint64_t need = padded(len);
        Chunk* c = &chunks_[cur_];
        const bool fits = !spilled_ && rec_at_ + need <= alloc_ && c->records.size() < max_records_;
        const bool full = !c->records.empty() && rec_at_ + need > target_;
        if (!fits || full) {
            // the record starts the next chunk
            if (!spilled_) {
                spill_.assign(c->h_seq + rec_at_, c->used - rec_at_);
                c->used = rec_at_;
            }
            int rc = flush();
            if (rc == CUSBF_OK) rc = take_next();
            if (rc == CUSBF_OK && need > alloc_) {
                // a record larger than the staging: finish the other chunk, then grow
                rc = finish(chunks_[cur_ ^ 1]);
                if (rc == CUSBF_OK) rc = allocate(need);
            }
            if (rc != CUSBF_OK) return rc;
            c = &chunks_[cur_];
            std::memcpy(c->h_seq, spill_.data(), spill_.size());
            c->used = spill_.size();
            rec_at_ = 0;
        }
        if (need > len) {
            std::memset(c->h_seq + c->used, sep_, need - len);
            c->used += need - len;
        }
        RecordMeta m{index, 0, 0, rec_at_, len};
        if (query_ && consumer_) {
            m.header_at = c->headers.size();
            m.header_len = header.size();
            c->headers.append(header);
        }
        c->records.push_back(m);
        ++records_;
        bases_ += len;
        return CUSBF_OK;
    }

    // (Re)allocates both staging sets and the device buffers for `bytes` of
    // sequence per chunk; both chunks must be idle.
    int allocate(uint64_t bytes) {
        (void)hipStreamSynchronize(stream_);
        free_buffers();
        alloc_ = (bytes + 63) / 64 * 64;
        max_records_ = std::max<uint64_t>(1024, alloc_ / kBytesPerRecordSlot);
        hipError_t e = hipSuccess;
        for (Chunk& c : chunks_) {
            if (e == hipSuccess) e = hipHostMalloc(&c.h_seq, alloc_, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc(&c.h_off, (max_records_ + 1) * 8, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc(&c.h_req, 2 * max_records_ * 8, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc(&c.h_counts, 2 * max_records_ * 8, hipHostMallocDefault);
            if (e == hipSuccess && query_) e = hipHostMalloc(&c.h_hits, alloc_, hipHostMallocDefault);
        }
        if (e == hipSuccess) e = hipMalloc(&d_seq_, alloc_);
        if (e == hipSuccess) e = hipMalloc(&d_off_, (max_records_ + 1) * 8);
        if (e == hipSuccess) e = hipMalloc(&d_req_, 2 * max_records_ * 8);
        if (e == hipSuccess) e = hipMalloc(&d_counts_, 2 * max_records_ * 8);
        if (e == hipSuccess && query_) e = hipMalloc(&d_hits_, alloc_);
        if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
        // Touch the staging the parser writes, and the hits a consumer reads, on
        // the host pool now rather than page by page later (first-touch faults);
        // the record tables are touched as records use them.
        for (Chunk& c : chunks_) {
            ci::touch_pages(ci::host_pool(), c.h_seq, alloc_);
            if (query_ && consumer_) ci::touch_pages(ci::host_pool(), c.h_hits, alloc_);
        }
        return CUSBF_OK;
    }

    void free_buffers() {
        for (Chunk& c : chunks_) {
            if (c.h_seq) (void)hipHostFree(c.h_seq);
            if (c.h_off) (void)hipHostFree(c.h_off);
            if (c.h_req) (void)hipHostFree(c.h_req);
            if (c.h_counts) (void)hipHostFree(c.h_counts);
            if (c.h_hits) (void)hipHostFree(c.h_hits);
            c.h_seq = nullptr;
            c.h_off = c.h_req = c.h_counts = nullptr;
            c.h_hits = nullptr;
        }
        if (d_seq_) (void)hipFree(d_seq_);
        if (d_off_) (void)hipFree(d_off_);
        if (d_req_) (void)hipFree(d_req_);
        if (d_counts_) (void)hipFree(d_counts_);
        if (d_hits_) (void)hipFree(d_hits_);
        d_seq_ = nullptr;
        d_off_ = d_req_ = d_counts_ = nullptr;
        d_hits_ = nullptr;
    }

    // Enqueues the current chunk's device work (nothing when it holds no record).
    int flush() {
        Chunk& c = chunks_[cur_];
        const uint64_t n = c.records.size();
        if (n == 0) return CUSBF_OK;
        const uint64_t symbols = c.used / width_;
        for (uint64_t r = 0; r < n; ++r) {
            const RecordMeta& m = c.records[r];
            const uint64_t first = m.byte_at / width_;
            const uint64_t syms = m.bytes / width_;
            c.h_off[r] = first;
            c.h_req[2 * r] = first;
            c.h_req[2 * r + 1] = syms >= k_ ? syms - k_ + 1 : 0;
        }
        c.h_off[n] = symbols;
        hipError_t e = hipMemcpyAsync(d_seq_, c.h_seq, c.used, hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(d_off_, c.h_off, (n + 1) * 8, hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(d_req_, c.h_req, 2 * n * 8, hipMemcpyHostToDevice, stream_);
        if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
        const uint64_t windows = symbols >= k_ ? symbols - k_ + 1 : 0;
        int rc = query_ ? cusbf_contains_batched_async(f_, d_seq_, symbols, d_off_, n, d_hits_, alloc_, stream_)
                        : cusbf_insert_batched_async(f_, d_seq_, symbols, d_off_, n, stream_);
        if (rc != CUSBF_OK) return rc;
        rc = ci::launch_request_counts(f_, d_seq_, c.used, query_ ? d_hits_ : nullptr, d_req_, n,
                                       d_counts_, stream_);
        if (rc != CUSBF_OK) return rc;
        e = hipMemcpyAsync(c.h_counts, d_counts_, 2 * n * 8, hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess && query_ && windows) {
            e = hipMemcpyAsync(c.h_hits, d_hits_, windows, hipMemcpyDeviceToHost, stream_);
        }
        if (e == hipSuccess) e = hipEventRecord(c.done, stream_);
        if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
        c.in_flight = true;
        return CUSBF_OK;
    }

    // Switches to the other staging set, finishing the chunk that used it.
    int take_next() {
        cur_ ^= 1;
        return finish(chunks_[cur_]);
    }

    // Waits for a chunk's device work, adds its counts to the report and hands
    // its records to the consumer in source order; leaves the chunk empty.
    int finish(Chunk& c) {
        if (c.in_flight) {
            const hipError_t e = hipEventSynchronize(c.done);
            if (e != hipSuccess) return ci::set_hip_error(e, __FILE__, __LINE__);
            c.in_flight = false;
            for (size_t r = 0; r < c.records.size(); ++r) {
                const RecordMeta& m = c.records[r];
                valid_ += c.h_counts[2 * r];
                positive_ += c.h_counts[2 * r + 1];
                if (query_ && consumer_) {
                    const uint64_t syms = m.bytes / width_;
                    const uint64_t hits = syms >= k_ ? syms - k_ + 1 : 0;
                    cusbf_fastx_record_view v{};
                    v.record_index = m.index;
                    v.header = c.headers.data() + m.header_at;
                    v.header_len = m.header_len;
                    v.sequence = c.h_seq + m.byte_at;
                    v.sequence_len = m.bytes;
                    v.queried_bases = m.bytes;
                    v.queried_kmers = c.h_counts[2 * r];
                    v.positive_kmers = c.h_counts[2 * r + 1];
                    v.hits = hits ? c.h_hits + m.byte_at / width_ : nullptr;
                    v.hits_len = hits;
                    if (!stopped_ && consumer_(&v, user_) != 0) stopped_ = true;
                }
            }
        }
        c.records.clear();
        c.headers.clear();
        c.used = 0;
        return CUSBF_OK;
    }

    // Finishes both chunks, oldest first.
    int drain() {
        int rc = finish(chunks_[cur_ ^ 1]);
        if (rc == CUSBF_OK) rc = finish(chunks_[cur_]);
        return rc;
    }

    cusbf_filter* f_;
    bool query_;
    hipStream_t stream_;
    cusbf_fastx_consumer consumer_;
    void* user_;
    std::string source_;
    uint32_t width_;
    uint32_t k_;
    char sep_;
    uint64_t target_ = 0, alloc_ = 0, max_records_ = 0;
    bool stopped_ = false;
    Chunk chunks_[2];
    char* d_seq_ = nullptr;  // device side of the chunk in flight (one stream orders them)
    uint64_t* d_off_ = nullptr;
    uint64_t* d_req_ = nullptr;
    uint64_t* d_counts_ = nullptr;
    uint8_t* d_hits_ = nullptr;
    int cur_ = 0;
    uint64_t rec_at_ = 0;
    bool spilled_ = false;
    std::string spill_;
    uint64_t records_ = 0, bases_ = 0, valid_ = 0, positive_ = 0;

   public:
    bool stopped() const { return stopped_; }
};

int run_pipeline(cusbf_filter* f, bool query, Source& src, const char* source_name,
                 double fill_fraction, void* stream, cusbf_fastx_consumer consumer, void* user,
                 cusbf_insert_report* ins, cusbf_query_report* qry) {
    Pipeline p(f, query, reinterpret_cast<hipStream_t>(stream), consumer, user, source_name);
    int rc = p.init(fill_fraction);
    if (rc == CUSBF_OK) rc = p.run(src);
    if (rc == CUSBF_OK) {
        if (ins) *ins = p.insert_report();
        if (qry) *qry = p.query_report();
    }
    return rc;
}

// Opens `path` as a source: gzip stream or mmap'd file.
int open_file(const char* path, std::unique_ptr<Source>& out) {
    if (!path) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null path");
    if (is_gzip(path)) {
        auto gz = std::make_unique<GzipSource>();
        if (!gz->open(path)) {
            return ci::set_error(CUSBF_ERR_IO, (std::string("Failed to open FASTA/FASTQ file: ") + path).c_str());
        }
        out = std::move(gz);
        return CUSBF_OK;
    }
    auto m = std::make_unique<MappedFile>();
    if (const char* step = m->open(path)) {
        return ci::set_error(CUSBF_ERR_IO,
                             (std::string("Failed to ") + step + " FASTA/FASTQ file: " + path).c_str());
    }
    out = std::move(m);
    return CUSBF_OK;
}

class ParseOnly final : public SequenceSink {
   public:
    void append(std::string_view piece) override { seq.append(piece.data(), piece.size()); }
    std::string seq;
};

}  // namespace

extern "C" {

int cusbf_insert_fastx_buffer(cusbf_filter* f, const char* text, uint64_t bytes,
                              const char* source_name, double fill_fraction, void* stream,
                              cusbf_insert_report* report) {
    if (!f || (!text && bytes)) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    MemorySource src(text, bytes);
    return run_pipeline(f, false, src, source_name ? source_name : "<stream>", fill_fraction, stream,
                        nullptr, nullptr, report, nullptr);
}

int cusbf_insert_fastx_reader(cusbf_filter* f, cusbf_read_fn read, void* user,
                              const char* source_name, double fill_fraction, void* stream,
                              cusbf_insert_report* report) {
    if (!f || !read) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    CallbackSource src(read, user);
    return run_pipeline(f, false, src, source_name ? source_name : "<stream>", fill_fraction, stream,
                        nullptr, nullptr, report, nullptr);
}

int cusbf_insert_fastx_file(cusbf_filter* f, const char* path, double fill_fraction, void* stream,
                            cusbf_insert_report* report) {
    if (!f) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    std::unique_ptr<Source> src;
    const int rc = open_file(path, src);
    if (rc != CUSBF_OK) return rc;
    return run_pipeline(f, false, *src, path, fill_fraction, stream, nullptr, nullptr, report, nullptr);
}

int cusbf_query_fastx_buffer(const cusbf_filter* f, const char* text, uint64_t bytes,
                             const char* source_name, double fill_fraction, void* stream,
                             cusbf_fastx_consumer consumer, void* user, cusbf_query_report* report) {
    if (!f || (!text && bytes)) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    MemorySource src(text, bytes);
    return run_pipeline(const_cast<cusbf_filter*>(f), true, src, source_name ? source_name : "<stream>",
                        fill_fraction, stream, consumer, user, nullptr, report);
}

int cusbf_query_fastx_reader(const cusbf_filter* f, cusbf_read_fn read, void* user_read,
                             const char* source_name, double fill_fraction, void* stream,
                             cusbf_fastx_consumer consumer, void* user, cusbf_query_report* report) {
    if (!f || !read) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    CallbackSource src(read, user_read);
    return run_pipeline(const_cast<cusbf_filter*>(f), true, src, source_name ? source_name : "<stream>",
                        fill_fraction, stream, consumer, user, nullptr, report);
}

int cusbf_query_fastx_file(const cusbf_filter* f, const char* path, double fill_fraction,
                           void* stream, cusbf_fastx_consumer consumer, void* user,
                           cusbf_query_report* report) {
    if (!f) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null filter");
    std::unique_ptr<Source> src;
    const int rc = open_file(path, src);
    if (rc != CUSBF_OK) return rc;
    return run_pipeline(const_cast<cusbf_filter*>(f), true, *src, path, fill_fraction, stream, consumer,
                        user, nullptr, report);
}

int cusbf_parse_fastx_buffer(const char* text, uint64_t bytes, const char* source_name,
                             cusbf_fastx_consumer consumer, void* user, uint64_t* records) {
    return cusbf_parse_fastx_buffer_chunked(text, bytes, source_name, 0, 0, consumer, user, records);
}

int cusbf_parse_fastx_buffer_chunked(const char* text, uint64_t bytes, const char* source_name,
                                     uint64_t chunk_bytes, uint64_t max_records,
                                     cusbf_fastx_consumer consumer, void* user, uint64_t* records) {
    if (!text && bytes) return ci::set_error(CUSBF_ERR_INVALID_ARGUMENT, "null argument");
    const std::string_view whole(text ? text : "", bytes);
    const uint64_t first = first_content_line(whole);
    if (first < whole.size() && whole[first] == '>') {
        // FASTA: the pipeline's split into chunks, without device work
        FastaSplitter sp(whole, 1);
        SplitBatch b;
        std::vector<char> out;
        uint64_t n = 0;
        for (uint64_t pos = first; pos < whole.size(); pos = b.cut) {
            sp.scan(pos, chunk_bytes ? chunk_bytes : kDefaultMaxChunk, max_records ? max_records : UINT64_MAX, &b);
            out.resize(std::max<uint64_t>(b.out_bytes, 1));
            sp.copy(b, out.data());
            for (const SplitRecord& r : b.records) {
                if (consumer) {
                    cusbf_fastx_record_view v{};
                    v.record_index = n;
                    v.header = whole.data() + r.header_at;
                    v.header_len = r.header_len;
                    v.sequence = out.data() + r.out_at;
                    v.sequence_len = r.bytes;
                    v.queried_bases = r.bytes;
                    if (consumer(&v, user) != 0) {
                        if (records) *records = n + 1;
                        return ci::set_error(CUSBF_ERR_IO, "FASTX record consumer stopped the query");
                    }
                }
                ++n;
            }
        }
        if (records) *records = n;
        return CUSBF_OK;
    }
    uint64_t n = 0, resume = 0;
    if (first < whole.size() && whole[first] == '@') {
        // FASTQ: split while the splitter can vouch for the chunks, the rest serially
        FastqSplitter sp(whole, 1);
        SplitBatch b;
        std::vector<char> out;
        uint64_t pos = first;
        while (pos < whole.size() &&
               sp.scan(pos, chunk_bytes ? chunk_bytes : kDefaultMaxChunk, max_records ? max_records : UINT64_MAX, &b)) {
            out.resize(std::max<uint64_t>(b.out_bytes, 1));
            sp.copy(b, out.data());
            for (const SplitRecord& r : b.records) {
                if (consumer) {
                    cusbf_fastx_record_view v{};
                    v.record_index = n;
                    v.header = whole.data() + r.header_at;
                    v.header_len = r.header_len;
                    v.sequence = out.data() + r.out_at;
                    v.sequence_len = r.bytes;
                    v.queried_bases = r.bytes;
                    if (consumer(&v, user) != 0) {
                        if (records) *records = n + 1;
                        return ci::set_error(CUSBF_ERR_IO, "FASTX record consumer stopped the query");
                    }
                }
                ++n;
            }
            pos = b.cut;
        }
        if (pos >= whole.size()) {
            if (records) *records = n;
            return CUSBF_OK;
        }
        resume = pos;
    }
    MemorySource src(text + resume, bytes - resume);
    Parser parser = resume ? Parser(src, count_newlines(text, resume), 2) : Parser(src);
    const std::string name = source_name ? source_name : "<stream>";
    std::string header;
    ParseOnly sink;
    ParseStatus st;
    for (;;) {
        sink.seq.clear();
        uint64_t len = 0;
        const int got = parser.next(header, sink, len, st);
        if (got < 0) {
            if (records) *records = n;
            if (st.code == CUSBF_ERR_IO) {
                return ci::set_error(CUSBF_ERR_IO, ("Failed to read FASTA/FASTQ input from " + name).c_str());
            }
            return ci::set_fastx_error(name.c_str(), st.line, st.column, st.detail);
        }
        if (got == 0) break;
        if (consumer) {
            cusbf_fastx_record_view v{};
            v.record_index = n;
            v.header = header.data();
            v.header_len = header.size();
            v.sequence = sink.seq.data();
            v.sequence_len = sink.seq.size();
            v.queried_bases = sink.seq.size();
            if (consumer(&v, user) != 0) {
                if (records) *records = n + 1;
                return ci::set_error(CUSBF_ERR_IO, "FASTX record consumer stopped the query");
            }
        }
        ++n;
    }
    if (records) *records = n;
    return CUSBF_OK;
}

}  // extern "C"
