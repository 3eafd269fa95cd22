// taas.hpp — the `.taas` v2 container (crates/SiMuse/crates/taas-format/src/
// types.rs, writer.rs, reader.rs, lib.rs:47-57):
//
//   b"TAAS2" | manifest length (u64 LE) | manifest (UTF-8 JSON) | payload
//
// manifest = {"version":"2.0.0","created":..,["game":..,]"sections":[..],
// "metadata":{..}} with each section {"name","kind","count","offset","size",
// "crc32"}: kind one of f32 i32 i64 u64 u8 str json blob, count = elements
// (bytes for str/json/blob), offset/size in the payload, crc32 = CRC-32
// (IEEE, reflected 0xEDB88320) of the section's bytes.  Keys are written in
// that order, without spaces, metadata keys sorted, as the reference's serde
// structs and BTreeMap serialise.  zstd-compressed sections (the reference's
// optional compression) are refused on reading: no zstd here.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace simuse {
namespace taas {

// CRC-32 of `n` bytes continuing from `prev` (the CRC of the bytes before
// them, 0 to start), as zlib's crc32(prev, data, n).
inline uint32_t crc32(const void* data, size_t n, uint32_t prev = 0) {
    struct Table {
        uint32_t t[256];
        Table() {
            for (uint32_t i = 0; i < 256; ++i) {
                uint32_t c = i;
                for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
                t[i] = c;
            }
        }
    };
    static const Table tab;  // initialised once, thread-safe
    const uint32_t* table = tab.t;
    const uint8_t* p = static_cast<const uint8_t*>(data);
    uint32_t c = prev ^ 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

inline size_t kind_size(const std::string& kind) {
    if (kind == "f32" || kind == "i32") return 4;
    if (kind == "i64" || kind == "u64") return 8;
    return 1;  // u8, str, json, blob
}

inline void json_string(std::string& out, const std::string& s) {
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", c);
                    out += buf;
                } else {
                    out += static_cast<char>(c);
                }
        }
    }
    out += '"';
}

struct Section {
    std::string name, kind;
    uint64_t count = 0, offset = 0, size = 0;
    uint32_t crc = 0;
    bool has_crc = false;
    bool compressed = false;
};

class Writer {
public:
    std::string game, created;
    std::map<std::string, std::string> metadata;

    void push(const std::string& name, const std::string& kind, uint64_t count, const void* data, size_t bytes) {
        Section s;
        s.name = name;
        s.kind = kind;
        s.count = count;
        s.offset = payload_.size();
        s.size = bytes;
        s.crc = crc32(data, bytes);
        s.has_crc = true;
        const uint8_t* p = static_cast<const uint8_t*>(data);
        payload_.insert(payload_.end(), p, p + bytes);
        sections_.push_back(s);
    }

    std::string manifest() const {
        std::string m = "{\"version\":";
        json_string(m, "2.0.0");
        m += ",\"created\":";
        json_string(m, created.empty() ? now_utc() : created);
        if (!game.empty()) {
            m += ",\"game\":";
            json_string(m, game);
        }
        m += ",\"sections\":[";
        for (size_t i = 0; i < sections_.size(); ++i) {
            const Section& s = sections_[i];
            if (i) m += ',';
            m += "{\"name\":";
            json_string(m, s.name);
            m += ",\"kind\":";
            json_string(m, s.kind);
            m += ",\"count\":" + std::to_string(s.count) + ",\"offset\":" + std::to_string(s.offset) +
                 ",\"size\":" + std::to_string(s.size) + ",\"crc32\":" + std::to_string(s.crc) + "}";
        }
        m += "],\"metadata\":{";
        bool first = true;
        for (const auto& kv : metadata) {
            if (!first) m += ',';
            first = false;
            json_string(m, kv.first);
            m += ':';
            json_string(m, kv.second);
        }
        m += "}}";
        return m;
    }

    // 0 on success.
    int to_file(const char* path) const {
        const std::string m = manifest();
        FILE* f = std::fopen(path, "wb");
        if (!f) return -1;
        const uint64_t len = m.size();
        uint8_t lb[8];
        for (int i = 0; i < 8; ++i) lb[i] = static_cast<uint8_t>(len >> (8 * i));
        bool ok = std::fwrite("TAAS2", 1, 5, f) == 5 && std::fwrite(lb, 1, 8, f) == 8 &&
                  std::fwrite(m.data(), 1, m.size(), f) == m.size() &&
                  (payload_.empty() || std::fwrite(payload_.data(), 1, payload_.size(), f) == payload_.size());
        ok = (std::fclose(f) == 0) && ok;
        return ok ? 0 : -1;
    }

    static std::string now_utc() {
        std::time_t t = std::time(nullptr);
        std::tm g{};
        gmtime_r(&t, &g);
        char buf[32];
        std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%SZ", &g);
        return buf;
    }

private:
    std::vector<Section> sections_;
    std::vector<uint8_t> payload_;
};

// Minimal JSON reader for the manifest: objects, arrays, strings, integers,
// true/false/null.  Values other than the ones the manifest uses are skipped.
class Json {
public:
    explicit Json(const std::string& s) : s_(s) {}

    bool manifest(std::vector<Section>& secs, std::map<std::string, std::string>& meta, std::string& version,
                  std::string& err) {
        if (!expect('{')) return fail(err, "manifest is not an object");
        if (peek() == '}') return ++i_, true;
        for (;;) {
            std::string key;
            if (!string(key) || !expect(':')) return fail(err, "bad manifest key");
            if (key == "sections") {
                if (!expect('[')) return fail(err, "sections is not an array");
                if (peek() == ']') {
                    ++i_;
                } else {
                    for (;;) {
                        Section s;
                        if (!section(s)) return fail(err, "bad section entry");
                        secs.push_back(s);
                        if (peek() == ',') { ++i_; continue; }
                        if (!expect(']')) return fail(err, "bad sections array");
                        break;
                    }
                }
            } else if (key == "metadata") {
                if (!expect('{')) return fail(err, "metadata is not an object");
                if (peek() == '}') {
                    ++i_;
                } else {
                    for (;;) {
                        std::string k, v;
                        if (!string(k) || !expect(':') || !string(v)) return fail(err, "bad metadata entry");
                        meta[k] = v;
                        if (peek() == ',') { ++i_; continue; }
                        if (!expect('}')) return fail(err, "bad metadata object");
                        break;
                    }
                }
            } else if (key == "version") {
                if (!string(version)) return fail(err, "bad version");
            } else if (!skip()) {
                return fail(err, "bad manifest value");
            }
            if (peek() == ',') { ++i_; continue; }
            if (!expect('}')) return fail(err, "bad manifest object");
            return true;
        }
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    static bool fail(std::string& err, const char* msg) {
        err = msg;
        return false;
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
    }
    char peek() {
        ws();
        return i_ < s_.size() ? s_[i_] : '\0';
    }
    bool expect(char c) {
        if (peek() != c) return false;
        ++i_;
        return true;
    }
    bool string(std::string& out) {
        if (!expect('"')) return false;
        out.clear();
        while (i_ < s_.size()) {
            char c = s_[i_++];
            if (c == '"') return true;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) return false;
            char e = s_[i_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned v;
                    if (!hex4(v)) return false;
                    if (v >= 0xD800 && v < 0xDC00) {  // high surrogate: a low one must follow
                        unsigned lo;
                        if (i_ + 2 > s_.size() || s_[i_] != '\\' || s_[i_ + 1] != 'u') return false;
                        i_ += 2;
                        if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return false;
                        v = 0x10000 + ((v - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (v >= 0xDC00 && v < 0xE000) {
                        return false;
                    }
                    utf8(out, v);
                    break;
                }
                default: return false;
            }
        }
        return false;
    }
    bool hex4(unsigned& v) {
        if (i_ + 4 > s_.size()) return false;
        v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s_[i_ + k];
            const int d = c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'f' ? c - 'a' + 10
                        : c >= 'A' && c <= 'F' ? c - 'A' + 10 : -1;
            if (d < 0) return false;
            v = v * 16 + static_cast<unsigned>(d);
        }
        i_ += 4;
        return true;
    }
    static void utf8(std::string& out, unsigned v) {
        if (v < 0x80) {
            out += static_cast<char>(v);
        } else if (v < 0x800) {
            out += static_cast<char>(0xC0 | (v >> 6));
            out += static_cast<char>(0x80 | (v & 0x3F));
        } else if (v < 0x10000) {
            out += static_cast<char>(0xE0 | (v >> 12));
            out += static_cast<char>(0x80 | ((v >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (v & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (v >> 18));
            out += static_cast<char>(0x80 | ((v >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((v >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (v & 0x3F));
        }
    }
    // A non-negative integer below 2^64.
    bool number(uint64_t& v) {
        ws();
        size_t j = i_;
        v = 0;
        while (j < s_.size() && s_[j] >= '0' && s_[j] <= '9') {
            const uint64_t d = static_cast<uint64_t>(s_[j] - '0');
            if (v > (~uint64_t(0) - d) / 10) return false;
            v = v * 10 + d;
            ++j;
        }
        if (j == i_) return false;
        i_ = j;
        return true;
    }
    bool skip() {  // any value
        const char c = peek();
        if (c == '"') {
            std::string t;
            return string(t);
        }
        if (c == '{' || c == '[') {
            const char close = c == '{' ? '}' : ']';
            ++i_;
            if (peek() == close) return ++i_, true;
            for (;;) {
                if (c == '{') {
                    std::string k;
                    if (!string(k) || !expect(':')) return false;
                }
                if (!skip()) return false;
                if (peek() == ',') { ++i_; continue; }
                return expect(close);
            }
        }
        // number, true, false, null
        const size_t j0 = i_;
        while (i_ < s_.size() && s_[i_] != ',' && s_[i_] != '}' && s_[i_] != ']' && s_[i_] != ' ' &&
               s_[i_] != '\n')
            ++i_;
        return i_ > j0;
    }
    bool section(Section& s) {
        if (!expect('{')) return false;
        if (peek() == '}') return ++i_, true;
        for (;;) {
            std::string key;
            if (!string(key) || !expect(':')) return false;
            bool ok = true;
            if (key == "name") ok = string(s.name);
            else if (key == "kind") ok = string(s.kind);
            else if (key == "count") ok = number(s.count);
            else if (key == "offset") ok = number(s.offset);
            else if (key == "size") ok = number(s.size);
            else if (key == "crc32") {
                uint64_t c = 0;
                ok = number(c);
                s.crc = static_cast<uint32_t>(c);
                s.has_crc = true;
            } else if (key == "compression") {
                if (peek() == 'n') ok = skip();
                else {
                    std::string t;
                    ok = string(t);
                    s.compressed = true;
                }
            } else {
                ok = skip();
            }
            if (!ok) return false;
            if (peek() == ',') { ++i_; continue; }
            return expect('}');
        }
    }
};

class Reader {
public:
    std::vector<Section> sections;
    std::map<std::string, std::string> metadata;
    std::string version, error;

    // false with `error` set on a bad file.
    bool open(const char* path) {
        FILE* f = std::fopen(path, "rb");
        if (!f) return err("cannot open file");
        std::fseek(f, 0, SEEK_END);
        const long len = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (len < 13) {
            std::fclose(f);
            return err("file too short");
        }
        try {
            bytes_.resize(static_cast<size_t>(len));
        } catch (const std::bad_alloc&) {
            std::fclose(f);
            return err("file too large for host memory");
        }
        const bool ok = std::fread(bytes_.data(), 1, bytes_.size(), f) == bytes_.size();
        std::fclose(f);
        if (!ok) return err("read failed");
        if (std::memcmp(bytes_.data(), "TAAS2", 5) != 0) return err("invalid magic (expected TAAS2)");
        uint64_t hl = 0;
        for (int i = 0; i < 8; ++i) hl |= static_cast<uint64_t>(bytes_[5 + i]) << (8 * i);
        if (hl > bytes_.size() - 13) return err("manifest truncated");
        const std::string man(reinterpret_cast<const char*>(bytes_.data() + 13), static_cast<size_t>(hl));
        data_ = 13 + static_cast<size_t>(hl);
        Json j(man);
        if (!j.manifest(sections, metadata, version, error)) return false;
        const uint64_t avail = bytes_.size() - data_;
        for (const Section& s : sections)
            if (s.offset > avail || s.size > avail - s.offset) return err("a section exceeds the file");
        return true;
    }

    const Section* find(const std::string& name) const {
        for (const Section& s : sections)
            if (s.name == name) return &s;
        return nullptr;
    }

    // Bytes of section `name` of `kind` holding exactly `count` elements,
    // CRC-checked; nullptr with `error` set otherwise.
    const uint8_t* get(const std::string& name, const std::string& kind, uint64_t count) {
        const Section* s = find(name);
        if (!s) return errp("missing section " + name);
        if (s->kind != kind) return errp("section " + name + " is " + s->kind + ", expected " + kind);
        if (s->compressed) return errp("section " + name + " is compressed (zstd is not supported)");
        if (s->count != count || s->size / kind_size(kind) != count || s->size % kind_size(kind))
            return errp("section " + name + " has " + std::to_string(s->count) + " elements, expected " +
                        std::to_string(count));
        const uint8_t* p = bytes_.data() + data_ + s->offset;
        if (s->has_crc && crc32(p, static_cast<size_t>(s->size)) != s->crc) return errp("CRC mismatch in " + name);
        return p;
    }

private:
    std::vector<uint8_t> bytes_;
    size_t data_ = 0;
    bool err(const char* m) {
        error = m;
        return false;
    }
    const uint8_t* errp(const std::string& m) {
        error = m;
        return nullptr;
    }
};

}  // namespace taas
}  // namespace simuse
