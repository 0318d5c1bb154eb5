// Parquet files of the observable and state tables (reference src/output.rs),
// written without Arrow: the host library has no third-party dependency.
//
// The files hold what the reference's ArrowWriter holds for the same rows --
// the same column names, order and types (Boolean, UInt64 as INT64 with the
// unsigned 64-bit integer annotation, Float64 as DOUBLE), all columns
// required -- and, like the reference's (output.rs:48,138), its pages are
// Snappy-compressed, so a reader sees the same table.  Values are PLAIN
// encoded (the reference's writer may also use dictionary pages; a reader
// decodes both the same).  Like the reference (output.rs:32-33, 101-113),
// rows go to `<path with extension "parquet.tmp">`, renamed to `path` on
// close.
//
// Format: "PAR1", then per row group and column one or more data pages
// (Thrift-compact PageHeader + the Snappy block of the PLAIN values; required
// top-level columns carry no repetition or definition levels), then the
// Thrift-compact FileMetaData, its length (4 bytes little-endian) and "PAR1"
// (Apache Parquet format specification, parquet.thrift).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "error.hpp"
#include "snappy.hpp"

namespace vegas {

namespace parquet_detail {

// Thrift compact protocol: field headers carry the id delta and the type in
// one byte; integers are zigzag varints.
class Compact {
  public:
    enum : uint8_t { kTrue = 1, kFalse = 2, kByte = 3, kI32 = 5, kI64 = 6, kBinary = 8,
                     kList = 9, kStruct = 12 };

    void i32(int16_t id, int32_t v) {
        field(id, kI32);
        varint(zigzag(v));
    }
    void i64(int16_t id, int64_t v) {
        field(id, kI64);
        varint(zigzag(v));
    }
    void i8(int16_t id, int8_t v) {
        field(id, kByte);
        out_.push_back((char)v);
    }
    void boolean(int16_t id, bool v) { field(id, v ? kTrue : kFalse); }
    void str(int16_t id, const std::string& s) {
        field(id, kBinary);
        bytes(s);
    }
    void begin_struct(int16_t id) {
        field(id, kStruct);
        push();
    }
    void end_struct() { pop(); }
    void begin_list(int16_t id, uint8_t elem, size_t n) {
        field(id, kList);
        if (n < 15) {
            out_.push_back((char)((n << 4) | elem));
        } else {
            out_.push_back((char)(0xF0 | elem));
            varint(n);
        }
    }
    // Elements of lists.
    void elem_i32(int32_t v) { varint(zigzag(v)); }
    void elem_str(const std::string& s) { bytes(s); }
    void begin_elem_struct() { push(); }
    void end_elem_struct() { pop(); }
    void stop() { out_.push_back(0); }  // end of the outermost struct

    const std::string& data() const { return out_; }

  private:
    static uint64_t zigzag(int64_t v) { return ((uint64_t)v << 1) ^ (uint64_t)(v >> 63); }
    void varint(uint64_t v) {
        while (v >= 0x80) {
            out_.push_back((char)((v & 0x7F) | 0x80));
            v >>= 7;
        }
        out_.push_back((char)v);
    }
    void bytes(const std::string& s) {
        varint(s.size());
        out_ += s;
    }
    void field(int16_t id, uint8_t type) {
        const int delta = id - last_;
        if (delta > 0 && delta <= 15) {
            out_.push_back((char)((delta << 4) | type));
        } else {
            out_.push_back((char)type);
            varint(zigzag(id));
        }
        last_ = id;
    }
    void push() {
        stack_.push_back(last_);
        last_ = 0;
    }
    void pop() {
        out_.push_back(0);
        last_ = stack_.back();
        stack_.pop_back();
    }
    std::string out_;
    std::vector<int16_t> stack_;
    int16_t last_ = 0;
};

// Enumerations of parquet.thrift used here.
enum : int32_t { kBoolean = 0, kInt64 = 2, kDouble = 5 };  // Type
enum : int32_t { kRequired = 0 };                          // FieldRepetitionType
enum : int32_t { kUint64 = 14 };                           // ConvertedType
enum : int32_t { kPlain = 0, kRle = 3 };                   // Encoding
enum : int32_t { kSnappy = 1 };                            // CompressionCodec
enum : int32_t { kDataPage = 0 };                          // PageType

// Rust's Path::with_extension(ext): the file name's extension (after its
// last '.', unless that dot starts the name) is replaced, or ext appended.
inline std::string with_extension(const std::string& path, const std::string& ext) {
    const size_t slash = path.find_last_of('/');
    const size_t name = slash == std::string::npos ? 0 : slash + 1;
    const size_t dot = path.find_last_of('.');
    const bool has_ext = dot != std::string::npos && dot > name;
    return (has_ext ? path.substr(0, dot) : path) + "." + ext;
}

}  // namespace parquet_detail

class ParquetWriter {
  public:
    enum class Kind { Boolean, UInt64, Float64 };
    struct Column {
        std::string name;
        Kind kind;
    };
    static constexpr size_t kPageValues = size_t(1) << 20;  // values per data page

    ParquetWriter(const std::string& path, std::vector<Column> schema)
        : path_(path), tmp_(parquet_detail::with_extension(path, "parquet.tmp")),
          schema_(std::move(schema)) {
        f_ = std::fopen(tmp_.c_str(), "wb");
        if (!f_) throw Error(ErrorKind::Io, "cannot create " + tmp_);
        put("PAR1", 4);
    }
    ParquetWriter(const ParquetWriter&) = delete;
    ParquetWriter& operator=(const ParquetWriter&) = delete;
    // As the reference's Drop (output.rs:101-113): close and rename, and
    // report a failure on stderr rather than throw.
    ~ParquetWriter() {
        if (!f_) return;
        try {
            close();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "error closing parquet writer: %s\n", e.what());
        }
    }

    // One row group of n rows: cols[c] points at n values of column c --
    // uint8_t 0/1 (Boolean), uint64_t (UInt64) or double (Float64).
    void write(size_t n, const std::vector<const void*>& cols) {
        if (!f_) throw Error(ErrorKind::Io, "Writer has been closed");
        if (cols.size() != schema_.size())
            throw Error(ErrorKind::InvalidArgument, "one pointer per column is needed");
        if (n == 0) return;
        RowGroup rg;
        rg.rows = n;
        for (size_t c = 0; c < schema_.size(); ++c) {
            Chunk ch;
            ch.offset = pos_;
            for (size_t p0 = 0; p0 < n; p0 += kPageValues) {
                const size_t m = std::min(kPageValues, n - p0);
                const std::string values = plain(schema_[c].kind, cols[c], p0, m);
                const std::string packed = snappy_compress(values.data(), values.size());
                parquet_detail::Compact h;
                h.i32(1, parquet_detail::kDataPage);
                h.i32(2, (int32_t)values.size());
                h.i32(3, (int32_t)packed.size());
                h.begin_struct(5);
                h.i32(1, (int32_t)m);
                h.i32(2, parquet_detail::kPlain);
                h.i32(3, parquet_detail::kRle);
                h.i32(4, parquet_detail::kRle);
                h.end_struct();
                h.stop();
                put(h.data().data(), h.data().size());
                put(packed.data(), packed.size());
                ch.bytes += h.data().size() + packed.size();
                ch.raw += h.data().size() + values.size();
            }
            rg.chunks.push_back(ch);
        }
        rows_ += n;
        groups_.push_back(rg);
    }

    // Footer, then rename into place.
    void close() {
        if (!f_) return;
        parquet_detail::Compact m;
        using parquet_detail::Compact;
        m.i32(1, 1);  // version
        m.begin_list(2, Compact::kStruct, schema_.size() + 1);
        m.begin_elem_struct();  // root
        m.str(4, "schema");
        m.i32(5, (int32_t)schema_.size());
        m.end_elem_struct();
        for (const Column& c : schema_) {
            m.begin_elem_struct();
            m.i32(1, physical(c.kind));
            m.i32(3, parquet_detail::kRequired);
            m.str(4, c.name);
            if (c.kind == Kind::UInt64) {
                m.i32(6, parquet_detail::kUint64);
                m.begin_struct(10);  // LogicalType
                m.begin_struct(10);  // INTEGER
                m.i8(1, 64);
                m.boolean(2, false);
                m.end_struct();
                m.end_struct();
            }
            m.end_elem_struct();
        }
        m.i64(3, (int64_t)rows_);
        m.begin_list(4, Compact::kStruct, groups_.size());
        for (const RowGroup& g : groups_) {
            m.begin_elem_struct();
            m.begin_list(1, Compact::kStruct, g.chunks.size());
            uint64_t total = 0;
            for (size_t c = 0; c < g.chunks.size(); ++c) {
                const Chunk& ch = g.chunks[c];
                total += ch.raw;  // total_byte_size: uncompressed
                m.begin_elem_struct();
                m.i64(2, (int64_t)ch.offset);
                m.begin_struct(3);  // ColumnMetaData
                m.i32(1, physical(schema_[c].kind));
                m.begin_list(2, Compact::kI32, 1);
                m.elem_i32(parquet_detail::kPlain);
                m.begin_list(3, Compact::kBinary, 1);
                m.elem_str(schema_[c].name);
                m.i32(4, parquet_detail::kSnappy);
                m.i64(5, (int64_t)g.rows);
                m.i64(6, (int64_t)ch.raw);    // total_uncompressed_size
                m.i64(7, (int64_t)ch.bytes);  // total_compressed_size
                m.i64(9, (int64_t)ch.offset);
                m.end_struct();
                m.end_elem_struct();
            }
            m.i64(2, (int64_t)total);
            m.i64(3, (int64_t)g.rows);
            m.end_elem_struct();
        }
        m.str(6, "vegas-hip host library");
        m.stop();
        put(m.data().data(), m.data().size());
        const uint32_t len = (uint32_t)m.data().size();
        const unsigned char le[4] = {(unsigned char)len, (unsigned char)(len >> 8),
                                     (unsigned char)(len >> 16), (unsigned char)(len >> 24)};
        put(le, 4);
        put("PAR1", 4);
        const bool ok = std::fclose(f_) == 0;
        f_ = nullptr;
        if (!ok) throw Error(ErrorKind::Io, "cannot write " + tmp_);
        if (std::rename(tmp_.c_str(), path_.c_str()) != 0)
            throw Error(ErrorKind::Io, "cannot rename " + tmp_ + " to " + path_);
    }

    const std::string& path() const { return path_; }
    const std::string& temp_path() const { return tmp_; }
    uint64_t rows() const { return rows_; }

  private:
    struct Chunk {
        uint64_t offset = 0, bytes = 0, raw = 0;  // bytes as stored, and before compression
    };
    struct RowGroup {
        uint64_t rows = 0;
        std::vector<Chunk> chunks;
    };

    static int32_t physical(Kind k) {
        return k == Kind::Boolean ? parquet_detail::kBoolean
                                  : (k == Kind::UInt64 ? parquet_detail::kInt64
                                                       : parquet_detail::kDouble);
    }

    // PLAIN values [p0, p0 + m): booleans bit-packed LSB first, 64-bit
    // values little-endian.
    static std::string plain(Kind k, const void* col, size_t p0, size_t m) {
        std::string s;
        if (k == Kind::Boolean) {
            const uint8_t* v = static_cast<const uint8_t*>(col) + p0;
            s.assign((m + 7) / 8, '\0');
            for (size_t i = 0; i < m; ++i)
                if (v[i]) s[i >> 3] = (char)(s[i >> 3] | (1 << (i & 7)));
        } else {
            const unsigned char* v = static_cast<const unsigned char*>(col) + 8 * p0;
            s.resize(8 * m);
            for (size_t i = 0; i < m; ++i)
                for (int b = 0; b < 8; ++b) s[8 * i + b] = (char)le_byte(v + 8 * i, b);
        }
        return s;
    }
    // Byte b (little-endian order) of the 8-byte value at p, on any host.
    static unsigned char le_byte(const unsigned char* p, int b) {
        uint64_t u;
        std::memcpy(&u, p, 8);
        return (unsigned char)(u >> (8 * b));
    }

    void put(const void* p, size_t n) {
        if (std::fwrite(p, 1, n, f_) != n) throw Error(ErrorKind::Io, "cannot write " + tmp_);
        pos_ += n;
    }

    std::string path_, tmp_;
    std::vector<Column> schema_;
    std::FILE* f_ = nullptr;
    uint64_t pos_ = 0, rows_ = 0;
    std::vector<RowGroup> groups_;
};

// The observable table of output.rs:37-46.
inline std::vector<ParquetWriter::Column> observable_schema() {
    using K = ParquetWriter::Kind;
    return {{"relax", K::Boolean},      {"stage", K::UInt64},  {"step", K::UInt64},
            {"n", K::UInt64},           {"temperature", K::Float64}, {"field", K::Float64},
            {"energy", K::Float64},     {"magnetization", K::Float64}};
}

// The state table of output.rs:126-136.
inline std::vector<ParquetWriter::Column> state_schema() {
    using K = ParquetWriter::Kind;
    return {{"relax", K::Boolean},   {"stage", K::UInt64}, {"step", K::UInt64},
            {"temperature", K::Float64}, {"field", K::Float64}, {"id", K::UInt64},
            {"sx", K::Float64},      {"sy", K::Float64},   {"sz", K::Float64}};
}

}  // namespace vegas
