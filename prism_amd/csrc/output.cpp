// output.cpp -- VTK XML unstructured-grid and collection writers.
//
// The layout follows the VTK XML file format (UnstructuredGrid, version 0.1, little endian):
// Piece -> PointData, CellData, Points, Cells (connectivity, offsets, types), and for the
// compressed mode an AppendedData section in raw encoding whose arrays are addressed by byte
// offsets.  Each compressed array is [n_blocks, block_size, last_block_size, c_size_0 ...
// c_size_{n-1}] as UInt64 followed by the zlib streams of its blocks.
#include "output.h"

#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>

#include "../../include/prism_hip.h"
#include "internal.h"

namespace prism_hip {

namespace {

constexpr size_t BLOCK = 32768;

struct File {
    FILE* f{nullptr};
    std::string path;
    explicit File(const std::string& p) : path(p) {
        f = std::fopen(p.c_str(), "wb");
        if (!f) throw Error(PRISM_ERR_RUNTIME, "cannot open " + p + " for writing");
        std::setvbuf(f, nullptr, _IOFBF, 1 << 20);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw Error(PRISM_ERR_RUNTIME, "write to " + path + " failed");
    }
    void put(const std::string& s) { put(s.data(), s.size()); }
    void close() {
        FILE* g = f;
        f = nullptr;
        if (std::fclose(g) != 0) throw Error(PRISM_ERR_RUNTIME, "closing " + path + " failed");
    }
};

// Text of an XML attribute value.
std::string xmlAttr(const std::string& v) {
    std::string o;
    for (char c : v) {
        switch (c) {
            case '&': o += "&amp;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            case '"': o += "&quot;"; break;
            default: o += c;
        }
    }
    return o;
}

// A DataArray's payload: its raw bytes and how to print one element.
struct Payload {
    const char* type;
    const void* data;
    size_t count, elem;
};

void putAscii(File& f, const Payload& p) {
    std::string line;
    line.reserve(1 << 16);
    char buf[40];
    for (size_t i = 0; i < p.count; ++i) {
        char* e;
        if (std::strcmp(p.type, "Float64") == 0) {
            e = std::to_chars(buf, buf + sizeof(buf), static_cast<const double*>(p.data)[i]).ptr;
        } else if (std::strcmp(p.type, "Int64") == 0) {
            e = std::to_chars(buf, buf + sizeof(buf), static_cast<const int64_t*>(p.data)[i]).ptr;
        } else {
            e = std::to_chars(buf, buf + sizeof(buf), (int)static_cast<const int8_t*>(p.data)[i]).ptr;
        }
        line.append(buf, e);
        line.push_back(' ');
        if (line.size() > (1 << 16) - 64) {
            f.put(line);
            line.clear();
        }
    }
    line.push_back('\n');
    f.put(line);
}

// zlib-compressed blocks of one array, header first.
std::vector<unsigned char> compressArray(const Payload& p) {
    const size_t bytes = p.count * p.elem;
    const size_t nb = (bytes + BLOCK - 1) / BLOCK;
    std::vector<uint64_t> head(3 + nb);
    head[0] = nb;
    head[1] = BLOCK;
    head[2] = bytes % BLOCK;
    std::vector<unsigned char> body;
    const unsigned char* src = static_cast<const unsigned char*>(p.data);
    std::vector<unsigned char> tmp(compressBound(BLOCK));
    for (size_t b = 0; b < nb; ++b) {
        size_t len = std::min(BLOCK, bytes - b * BLOCK);
        uLongf out = (uLongf)tmp.size();
        if (compress2(tmp.data(), &out, src + b * BLOCK, (uLong)len, Z_DEFAULT_COMPRESSION) != Z_OK)
            throw Error(PRISM_ERR_RUNTIME, "zlib compression failed");
        head[3 + b] = out;
        body.insert(body.end(), tmp.begin(), tmp.begin() + out);
    }
    std::vector<unsigned char> all(head.size() * sizeof(uint64_t));
    std::memcpy(all.data(), head.data(), all.size());
    all.insert(all.end(), body.begin(), body.end());
    return all;
}

}  // namespace

void writeVtu(const std::string& path, int mode, const std::vector<double>& pts, const std::vector<int64_t>& off,
              const std::vector<int64_t>& conn, const std::vector<VtuArray>& arrays) {
    if (mode != 0 && mode != 1) throw Error(PRISM_ERR_INVALID, "unsupported write mode " + std::to_string(mode));
    const size_t nv = pts.size() / 3, nc = off.size() - 1;
    std::vector<int64_t> offsets(nc);
    std::vector<int8_t> types(nc);
    for (size_t c = 0; c < nc; ++c) {
        offsets[c] = off[c + 1];
        switch (off[c + 1] - off[c]) {
            case 8: types[c] = 12; break;
            case 4: types[c] = 10; break;
            case 6: types[c] = 13; break;
            default: throw Error(PRISM_ERR_RUNTIME, "Unsupported cell type for vtu export.");
        }
    }
    struct Item {
        std::string open;  // DataArray attributes before format
        Payload p;
    };
    std::vector<Item> items;
    auto dataAttr = [](const std::string& type, const std::string& name, int ncomp) {
        std::string s = "<DataArray type=\"" + type + "\"";
        if (!name.empty()) s += " Name=\"" + xmlAttr(name) + "\"";
        if (ncomp > 0) s += " NumberOfComponents=\"" + std::to_string(ncomp) + "\"";
        return s;
    };
    for (int pass = 0; pass < 2; ++pass)
        for (const auto& a : arrays)
            if (a.point == (pass == 0))
                items.push_back({dataAttr("Float64", a.name, a.ncomp), {"Float64", a.data.data(), a.data.size(), 8}});
    items.push_back({dataAttr("Float64", "", 3), {"Float64", pts.data(), pts.size(), 8}});
    items.push_back({dataAttr("Int64", "connectivity", 0), {"Int64", conn.data(), conn.size(), 8}});
    items.push_back({dataAttr("Int64", "offsets", 0), {"Int64", offsets.data(), offsets.size(), 8}});
    items.push_back({dataAttr("Int8", "types", 0), {"Int8", types.data(), types.size(), 1}});
    size_t n_point = 0, n_cell = 0;
    for (const auto& a : arrays) (a.point ? n_point : n_cell)++;

    std::vector<std::vector<unsigned char>> blobs;
    std::vector<size_t> where;
    if (mode == 1) {
        size_t at = 0;
        for (const auto& it : items) {
            blobs.push_back(compressArray(it.p));
            where.push_back(at);
            at += blobs.back().size();
        }
    }
    File f(path);
    f.put("<?xml version=\"1.0\"?>\n");
    f.put("<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\"");
    f.put(mode == 1 ? " compressor=\"vtkZLibDataCompressor\">\n" : ">\n");
    f.put("<UnstructuredGrid>\n");
    f.put("<Piece NumberOfPoints=\"" + std::to_string(nv) + "\" NumberOfCells=\"" + std::to_string(nc) + "\">\n");
    size_t k = 0;
    auto section = [&](const char* tag, size_t count) {
        f.put(std::string("<") + tag + ">\n");
        for (size_t i = 0; i < count; ++i, ++k) {
            if (mode == 1) {
                f.put(items[k].open + " format=\"appended\" offset=\"" + std::to_string(where[k]) + "\"/>\n");
            } else {
                f.put(items[k].open + " format=\"ascii\">\n");
                putAscii(f, items[k].p);
                f.put("</DataArray>\n");
            }
        }
        f.put(std::string("</") + tag + ">\n");
    };
    section("PointData", n_point);
    section("CellData", n_cell);
    section("Points", 1);
    section("Cells", 3);
    f.put("</Piece>\n</UnstructuredGrid>\n");
    if (mode == 1) {
        f.put("<AppendedData encoding=\"raw\">\n_");
        for (const auto& b : blobs) f.put(b.data(), b.size());
        f.put("\n</AppendedData>\n");
    }
    f.put("</VTKFile>\n");
    f.close();
}

void writePvd(const std::string& path, int64_t n, const double* times, const char* const* files) {
    File f(path);
    f.put("<?xml version=\"1.0\"?>\n");
    f.put("<VTKFile type=\"Collection\" version=\"0.1\">\n");
    f.put("  <Collection>\n");
    char buf[64];
    for (int64_t i = 0; i < n; ++i) {
        std::snprintf(buf, sizeof(buf), "%.6g", times[i]);
        f.put(std::string("    <DataSet timestep=\"") + buf + "\" file=\"" + xmlAttr(files[i]) + "\"/>\n");
    }
    f.put("  </Collection>\n");
    f.put("</VTKFile>\n");
    f.close();
}

}  // namespace prism_hip
