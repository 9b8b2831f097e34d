// gpart_verify_csr: independent exact checker of a partition of an int64 CSR.
// Weighted extension of the reference checker (tools/verify_csr.cpp:37-110):
// optional uint32 weight files, exact uint64 sums, arbitrary imbalance, and
// the same JSON keys.  Does not link the partitioner.
// Usage: indptr indices k partition binary|text [--imbalance X] [--edge-weights f] [--vertex-weights f] [--symmetric]
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

template <typename T>
static std::vector<T> read_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) throw std::runtime_error("cannot open " + path);
    const auto bytes = in.tellg();
    if (bytes < 0 || bytes % sizeof(T)) throw std::runtime_error("invalid file size " + path);
    std::vector<T> v(static_cast<std::size_t>(bytes) / sizeof(T));
    in.seekg(0);
    in.read(reinterpret_cast<char*>(v.data()), bytes);
    if (!in && bytes) throw std::runtime_error("cannot read " + path);
    return v;
}

int main(int argc, char** argv) {
    try {
        if (argc < 6) throw std::runtime_error("usage: indptr indices k partition binary|text [--imbalance X] [--edge-weights f] [--vertex-weights f] [--symmetric]");
        const int k = std::stoi(argv[3]);
        if (k < 2 || k > 128) throw std::runtime_error("invalid k");
        double imbalance = 1.03;
        bool symmetric = false;
        std::string ewf, vwf;
        for (int i = 6; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--symmetric") symmetric = true;
            else if (a == "--imbalance" && i + 1 < argc) imbalance = std::stod(argv[++i]);
            else if (a == "--edge-weights" && i + 1 < argc) ewf = argv[++i];
            else if (a == "--vertex-weights" && i + 1 < argc) vwf = argv[++i];
            else throw std::runtime_error("invalid flag " + a);
        }
        const auto off = read_file<std::int64_t>(argv[1]);
        const auto idx = read_file<std::int64_t>(argv[2]);
        if (off.empty() || off[0] != 0 || off.back() != static_cast<std::int64_t>(idx.size()))
            throw std::runtime_error("invalid CSR endpoints");
        const std::int64_t n = static_cast<std::int64_t>(off.size()) - 1;
        for (std::int64_t v = 0; v < n; ++v)
            if (off[v] > off[v + 1]) throw std::runtime_error("nonmonotone offsets");
        std::vector<std::uint32_t> ew, vw;
        if (!ewf.empty()) {
            ew = read_file<std::uint32_t>(ewf);
            if (ew.size() != idx.size()) throw std::runtime_error("edge weight length mismatch");
        }
        if (!vwf.empty()) {
            vw = read_file<std::uint32_t>(vwf);
            if (vw.size() != static_cast<std::size_t>(n)) throw std::runtime_error("vertex weight length mismatch");
        }
        std::vector<std::int32_t> labels;
        if (std::string(argv[5]) == "binary") {
            labels = read_file<std::int32_t>(argv[4]);
        } else if (std::string(argv[5]) == "text") {
            std::ifstream in(argv[4]);
            std::int64_t x;
            while (in >> x) labels.push_back(static_cast<std::int32_t>(x));
        } else
            throw std::runtime_error("invalid partition format");
        if (labels.size() != static_cast<std::size_t>(n)) throw std::runtime_error("partition length mismatch");
        std::vector<std::uint64_t> loads(static_cast<std::size_t>(k), 0);
        std::uint64_t total = 0;
        for (std::int64_t v = 0; v < n; ++v) {
            const auto p = labels[static_cast<std::size_t>(v)];
            if (p < 0 || p >= k) throw std::runtime_error("invalid label");
            const std::uint64_t w = vw.empty() ? 1 : vw[static_cast<std::size_t>(v)];
            loads[static_cast<std::size_t>(p)] += w;
            total += w;
        }
        std::uint64_t cut = 0, nonloop = 0;
        for (std::int64_t v = 0; v < n; ++v)
            for (std::int64_t e = off[v]; e < off[v + 1]; ++e) {
                const auto u = idx[static_cast<std::size_t>(e)];
                if (u < 0 || u >= n) throw std::runtime_error("invalid neighbor id");
                const std::uint64_t w = ew.empty() ? 1 : ew[static_cast<std::size_t>(e)];
                if (u != v) nonloop += w;
                if (labels[static_cast<std::size_t>(v)] != labels[static_cast<std::size_t>(u)]) cut += w;
            }
        if (symmetric)
            for (std::int64_t v = 0; v < n; ++v)
                for (std::int64_t e = off[v]; e < off[v + 1]; ++e) {
                    const auto u = idx[static_cast<std::size_t>(e)];
                    if (u == v) continue;
                    const auto b = idx.begin() + off[u], en = idx.begin() + off[u + 1];
                    const auto it = std::lower_bound(b, en, v);
                    if (it == en || *it != v) throw std::runtime_error("asymmetric edge");
                }
        if (cut % 2 || nonloop % 2) throw std::runtime_error("odd undirected edge/cut count");
        cut /= 2;
        nonloop /= 2;
        const long double raw = (long double)total * imbalance / k;
        const std::uint64_t capacity = (std::uint64_t)std::floor(raw);
        const std::uint64_t maximum = *std::max_element(loads.begin(), loads.end());
        std::cout.precision(17);
        std::cout << "{\"vertices\":" << n << ",\"undirected_edges\":" << nonloop << ",\"cut\":" << cut
                  << ",\"cut_ratio\":" << (nonloop ? (double)cut / (double)nonloop : 0) << ",\"loads\":[";
        for (int p = 0; p < k; ++p) std::cout << (p ? "," : "") << loads[static_cast<std::size_t>(p)];
        std::cout << "],\"max_load\":" << maximum << ",\"max_load_ratio\":" << (total ? (double)maximum * k / (double)total : 0)
                  << ",\"capacity\":" << capacity << ",\"capacity_ok\":" << (maximum <= capacity ? "true" : "false") << "}\n";
        return maximum <= capacity ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << e.what() << '\n';
        return 2;
    }
}
