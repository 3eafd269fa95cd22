// Command-line driver of the library's `.taas` v2 writer and reader
// (dioma-core_amd/csrc/taas.hpp) so that the CPU tests can compare them with
// the Python implementation byte for byte.
//
//   taas_check crc <text>     CRC-32 of the text's bytes, decimal
//   taas_check write <path>   the fixed sample file (see sample())
//   taas_check read <path>    one line per section: name kind count size ok|error
//                             then "meta <key> <value>" lines (JSON strings); exit 1 on a bad file
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "taas.hpp"

using simuse::taas::Reader;
using simuse::taas::Writer;

static void sample(Writer& w) {
    w.game = "dioma";
    w.created = "2026-01-01T00:00:00Z";
    w.metadata["world"] = "meadow";
    w.metadata["a\"quote"] = "tab\tnew\nline";
    const float f[3] = {1.0f, 2.0f, 3.0f};
    const int32_t i[2] = {10, 20};
    const int64_t l[1] = {int64_t(1) << 40};
    const uint64_t u[1] = {42};
    const uint8_t b[3] = {0, 255, 7};
    const char* s = "creature-7";
    const char* j = "{\"mut\":0.01,\"tags\":[\"a\"]}";
    const uint8_t raw[3] = {9, 8, 7};
    w.push("weights", "f32", 3, f, sizeof f);
    w.push("ids", "i32", 2, i, sizeof i);
    w.push("big", "i64", 1, l, sizeof l);
    w.push("stamps", "u64", 1, u, sizeof u);
    w.push("mask", "u8", 3, b, sizeof b);
    w.push("name", "str", std::strlen(s), s, std::strlen(s));
    w.push("genome", "json", std::strlen(j), j, std::strlen(j));
    w.push("raw", "blob", 3, raw, sizeof raw);
    w.push("empty", "f32", 0, nullptr, 0);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: taas_check crc|write|read <arg>\n");
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "crc") {
        std::printf("%u\n", simuse::taas::crc32(argv[2], std::strlen(argv[2])));
        return 0;
    }
    if (cmd == "write") {
        Writer w;
        sample(w);
        return w.to_file(argv[2]) == 0 ? 0 : 1;
    }
    if (cmd == "read") {
        Reader r;
        if (!r.open(argv[2])) {
            std::printf("error %s\n", r.error.c_str());
            return 1;
        }
        std::printf("version %s\n", r.version.c_str());
        for (const auto& s : r.sections) {
            const uint8_t* p = r.get(s.name, s.kind, s.count);
            std::printf("%s %s %llu %llu %s\n", s.name.c_str(), s.kind.c_str(),
                        static_cast<unsigned long long>(s.count), static_cast<unsigned long long>(s.size),
                        p ? "ok" : r.error.c_str());
        }
        for (const auto& kv : r.metadata) {
            std::string k, v;
            simuse::taas::json_string(k, kv.first);
            simuse::taas::json_string(v, kv.second);
            std::printf("meta %s %s\n", k.c_str(), v.c_str());
        }
        return 0;
    }
    return 2;
}
