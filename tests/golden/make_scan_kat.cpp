// tests/golden/make_scan_kat.cpp — writes scan_kat.json: the known-answer sets of the reference's scan tests
// (tests/test_device_scan.cpp:57-87: the pattern (i * 2654435761 + 7) mod 11 at the listed sizes, 700 zeros, a single 2^30
// value; tests/test_shapes.cpp:656-682: counts {2, 0, 3, 1} -> offsets {0, 2, 2, 5, 6}, total 6), with the expected
// outputs computed the way the reference's test computes them: std::exclusive_scan plus std::accumulate. Large outputs
// are stored as an FNV-1a checksum over their little-endian bytes plus the total and a few probes.
//   g++ -std=c++20 -O2 make_scan_kat.cpp -o /tmp/make_scan_kat && /tmp/make_scan_kat > scan_kat.json
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

static std::vector<std::uint32_t> pattern(std::size_t n) {
    std::vector<std::uint32_t> v(n);
    for (std::size_t i = 0; i < n; ++i) v[i] = static_cast<std::uint32_t>((i * 2654435761u + 7u) % 11u);
    return v;
}

static std::vector<std::uint32_t> expected(const std::vector<std::uint32_t>& v) {
    std::vector<std::uint32_t> out(v.size() + 1, 0);
    std::exclusive_scan(v.begin(), v.end(), out.begin(), std::uint32_t{0});
    out[v.size()] = std::accumulate(v.begin(), v.end(), std::uint32_t{0});
    return out;
}

static std::uint64_t fnv1a(const std::vector<std::uint32_t>& v) {
    std::uint64_t h = 1469598103934665603ull;
    for (std::uint32_t x : v)
        for (int b = 0; b < 4; ++b) {
            h ^= (x >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    return h;
}

static std::string list(const std::vector<std::uint32_t>& v) {
    std::string s = "[";
    for (std::size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
    return s + "]";
}

static void emit(const char* name, const char* kind, const std::vector<std::uint32_t>& in, bool first) {
    const std::vector<std::uint32_t> out = expected(in);
    std::printf("%s{\"name\": \"%s\", \"input\": \"%s\", \"n\": %zu, ", first ? "" : ",\n ", name, kind, in.size());
    if (in.size() <= 5000) {
        std::printf("\"values\": %s, \"expected\": %s}", list(in).c_str(), list(out).c_str());
    } else {
        std::printf("\"expected_fnv1a\": \"%llu\", \"expected_total\": %u, \"probes\": [[0, %u], [%zu, %u], [%zu, %u]]}",
                    static_cast<unsigned long long>(fnv1a(out)), out.back(), out[0], in.size() / 2, out[in.size() / 2], in.size() - 1,
                    out[in.size() - 1]);
    }
}

int main() {
    std::printf("{\"source\": \"tests/test_device_scan.cpp:57-87, tests/test_shapes.cpp:656-682 of the reference; std::exclusive_scan\",\n"
                " \"cases\": [");
    bool first = true;
    for (std::size_t n : {0, 1, 2, 37, 255, 256, 257, 1000, 4096 + 3, 256 * 256 + 17}) {
        emit(("pattern_" + std::to_string(n)).c_str(), "pattern", pattern(n), first);
        first = false;
    }
    emit("zeros_700", "values", std::vector<std::uint32_t>(700, 0), false);
    std::vector<std::uint32_t> big(700, 0);
    big[699] = 1u << 30;
    emit("spike_700", "values", big, false);
    emit("counts_from_2031", "values", {2, 0, 3, 1}, false);
    std::printf("]}\n");
}
