// Host I/O around the GPU path: the BABELSIM_MESH 3 text format
// (src/io/mesh_reader.cpp:44-111), per-rank result CSVs and metadata
// (src/parallel/parallel_writer.cpp:14-171) and nonuniform cell values for
// restarts (src/io/field_reader.cpp:143-241).  Cell ids in files are the
// caller's (original) ids; the device renumbering never reaches a file.
#include "io_host.hpp"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>

namespace bsim {
namespace fs = std::filesystem;

namespace {

[[noreturn]] void invalidMesh(const std::string& path, const std::string& what) {
    throw std::runtime_error("invalid BabelSim mesh file " + path + ": " + what);
}

template <class T> T take(std::istream& in, const std::string& path, const char* what) {
    T v{};
    if (!(in >> v)) invalidMesh(path, std::string("missing ") + what);
    return v;
}

int positiveCount(std::istream& in, const std::string& path, const char* what) {
    const long long v = take<long long>(in, path, what);
    if (v <= 0 || v >= std::numeric_limits<int>::max()) invalidMesh(path, std::string(what) + " must be positive");
    return static_cast<int>(v);
}

void keyword(std::istream& in, const std::string& path, const char* word) {
    if (take<std::string>(in, path, word) != word) invalidMesh(path, std::string("expected ") + word);
}

const char* kindNames[] = {"generic", "wall", "inlet", "outlet", "symmetry", "processor"};

int kindOf(const std::string& name, const std::string& path) {
    for (int k = 0; k < 6; ++k)
        if (name == kindNames[k]) return k;
    if (name == "mirror") return kPatchSymmetry;
    invalidMesh(path, "unknown patch kind " + name);
}

std::string rankDirectory(const std::string& timeDir, int rank) {
    std::ostringstream name;
    name << "rank-" << std::setfill('0') << std::setw(4) << rank;
    return (fs::path(timeDir) / name.str()).string();
}

void makeDirectories(const std::string& dir) {
    std::error_code code;
    fs::create_directories(dir, code);
    if (code) throw std::runtime_error("cannot create result directory: " + dir);
}

}  // namespace

MeshDesc readMeshFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open BabelSim mesh file: " + path);
    const std::string magic = take<std::string>(in, path, "file header");
    const int version = take<int>(in, path, "format version");
    if (magic != "BABELSIM_MESH" || version != 3) invalidMesh(path, "only BABELSIM_MESH version 3 is supported");
    MeshDesc d;
    keyword(in, path, "vertices");
    const int nv = positiveCount(in, path, "vertex count");
    d.vert.resize(3 * static_cast<std::size_t>(nv));
    for (double& x : d.vert) {
        x = take<double>(in, path, "vertex coordinate");
        if (!std::isfinite(x)) invalidMesh(path, "vertex must be finite");
    }
    keyword(in, path, "faces");
    const int nf = positiveCount(in, path, "face count");
    d.fpo.assign(1, 0);
    d.own.reserve(nf);
    d.nei.reserve(nf);
    d.patch.reserve(nf);
    for (int f = 0; f < nf; ++f) {
        if (take<std::string>(in, path, "face keyword") != "face") invalidMesh(path, "expected face");
        const int k = positiveCount(in, path, "face vertex count");
        d.own.push_back(take<int>(in, path, "face owner"));
        d.nei.push_back(take<int>(in, path, "face neighbour"));
        d.patch.push_back(take<int>(in, path, "face patch"));
        for (int j = 0; j < k; ++j) {
            const int v = take<int>(in, path, "face vertex index");
            if (v < 0 || v >= nv) invalidMesh(path, "face vertex index out of range");
            d.fpi.push_back(v);
        }
        d.fpo.push_back(static_cast<int>(d.fpi.size()));
    }
    keyword(in, path, "patches");
    const int np = positiveCount(in, path, "patch count");
    for (int p = 0; p < np; ++p) {
        if (take<std::string>(in, path, "patch keyword") != "patch") invalidMesh(path, "expected patch");
        const std::string name = take<std::string>(in, path, "patch name");
        const int kind = kindOf(take<std::string>(in, path, "patch kind"), path);
        d.patchName.push_back(name);
        d.patchKind.push_back(kind);
    }
    d.npatch = np;
    if (take<std::string>(in, path, "end marker") != "end") invalidMesh(path, "expected end");
    std::string trailing;
    if (in >> trailing) invalidMesh(path, "unexpected trailing token " + trailing);
    // Topology checks of Mesh::polyhedral (src/core/mesh.cpp:321-367).
    const std::string bad = d.topologyError();
    if (!bad.empty()) invalidMesh(path, bad);
    return d;
}

void writeMeshFile(const MeshDesc& d, const std::string& path) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot create mesh file: " + path);
    out << "BABELSIM_MESH 3\nvertices\n" << d.vert.size() / 3 << '\n' << std::setprecision(17);
    for (std::size_t v = 0; v < d.vert.size(); v += 3) out << d.vert[v] << ' ' << d.vert[v + 1] << ' ' << d.vert[v + 2] << '\n';
    out << "faces\n" << d.nfaces() << '\n';
    for (int f = 0; f < d.nfaces(); ++f) {
        out << "face " << d.fpo[f + 1] - d.fpo[f] << ' ' << d.own[f] << ' ' << d.nei[f] << ' ' << d.patch[f];
        for (int k = d.fpo[f]; k < d.fpo[f + 1]; ++k) out << ' ' << d.fpi[k];
        out << '\n';
    }
    out << "patches\n" << d.npatch << '\n';
    for (int p = 0; p < d.npatch; ++p) {
        const std::string name = p < static_cast<int>(d.patchName.size()) ? d.patchName[p] : "patch" + std::to_string(p);
        const int kind = p < static_cast<int>(d.patchKind.size()) ? d.patchKind[p] : kPatchGeneric;
        out << "patch\n" << name << '\n' << kindNames[kind] << '\n';
    }
    out << "end\n";
    if (!out) throw std::runtime_error("cannot write mesh file: " + path);
}

void writeFieldCsv(const std::string& timeDir, int rank, const std::string& name, int ncells, const int* globalIds,
                   const double* centres, int ncomp, const double* values) {
    if (name.empty()) throw std::invalid_argument("result output requires a named cell field");
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) throw std::invalid_argument("field components must be 1, 3 or 9");
    const std::string dir = rankDirectory(timeDir, rank);
    makeDirectories(dir);
    const std::string path = (fs::path(dir) / (name + ".csv")).string();
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot create field output: " + path);
    out << "global_id,x,y,z";
    for (int c = 0; c < ncomp; ++c) out << ",value" << c;
    out << '\n' << std::setprecision(17);
    for (int i = 0; i < ncells; ++i) {
        out << (globalIds ? globalIds[i] : i) << ',' << centres[3 * static_cast<long long>(i)] << ','
            << centres[3 * static_cast<long long>(i) + 1] << ',' << centres[3 * static_cast<long long>(i) + 2];
        for (int c = 0; c < ncomp; ++c) out << ',' << values[static_cast<long long>(ncomp) * i + c];
        out << '\n';
    }
    out.close();
    if (!out) throw std::runtime_error("cannot flush field output: " + path);
}

void writeResultMetadata(const std::string& timeDir, int rank, int ranks, const MeshDesc& d, const int* globalIds,
                         long long globalCells, const std::string& timeName, const std::vector<std::string>& names,
                         const std::vector<std::string>& types, const std::vector<int>& locations) {
    if (timeName.empty() || names.empty()) throw std::invalid_argument("result metadata needs a time name and fields");
    const std::string dir = rankDirectory(timeDir, rank);
    makeDirectories(dir);
    const int nc = d.ncells();
    {
        // Per cell: id, vertex count and the sorted distinct vertices of its
        // faces (the cell vertex lists of src/core/mesh.cpp:473-484).
        std::vector<std::vector<int>> cellFaces(nc);
        for (int f = 0; f < d.nfaces(); ++f) {
            cellFaces[d.own[f]].push_back(f);
            if (d.nei[f] >= 0) cellFaces[d.nei[f]].push_back(f);
        }
        const std::string path = (fs::path(dir) / "mesh.geometry").string();
        std::ofstream geo(path);
        if (!geo) throw std::runtime_error("cannot create result mesh provenance");
        geo << std::setprecision(17);
        for (int c = 0; c < nc; ++c) {
            std::set<int> unique;
            for (int f : cellFaces[c])
                for (int k = d.fpo[f]; k < d.fpo[f + 1]; ++k) unique.insert(d.fpi[k]);
            geo << (globalIds ? globalIds[c] : c) << ',' << unique.size();
            for (int v : unique) geo << ',' << d.vert[3 * v] << ',' << d.vert[3 * v + 1] << ',' << d.vert[3 * v + 2];
            geo << '\n';
        }
        geo.close();
        if (!geo) throw std::runtime_error("cannot flush result mesh provenance");
    }
    static const char* locationNames[] = {"cell", "face", "vertex"};
    const std::string path = (fs::path(dir) / "metadata.bs").string();
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot create result metadata: " + path);
    out << "format babelsim_result 3\n"
        << "time " << timeName << '\n'
        << "rank " << rank << '\n'
        << "ranks " << ranks << '\n'
        << "global_cell_count " << globalCells << '\n'
        << "owned_cells " << nc << '\n';
    for (std::size_t k = 0; k < names.size(); ++k) {
        if (names[k].empty() || types[k].empty()) throw std::invalid_argument("result metadata field is incomplete");
        if (locations[k] < 0 || locations[k] > 2) throw std::invalid_argument("unknown field location");
        out << "field " << names[k] << ' ' << types[k] << ' ' << locationNames[locations[k]] << '\n';
    }
    out.close();
    if (!out) throw std::runtime_error("cannot flush result metadata: " + path);
}

void readFieldValues(const std::string& path, const std::string& name, int ncells, int ncomp, double* out) {
    if (ncells <= 0) throw std::runtime_error("nonuniform field requires a mesh with global cell IDs");
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) throw std::invalid_argument("field components must be 1, 3 or 9");
    if (!fs::exists(path)) throw std::runtime_error("cannot open nonuniform field data: " + path);
    std::vector<double> values(static_cast<std::size_t>(ncells) * ncomp);
    std::vector<char> seen(ncells, 0);
    auto store = [&](long long id, const double* v, const std::string& source, std::size_t line) {
        if (id < 0 || id >= ncells)
            throw std::runtime_error("invalid global cell ID in " + source + ":" + std::to_string(line));
        if (seen[id]) throw std::runtime_error("duplicate global cell ID in " + source + ":" + std::to_string(line));
        std::copy(v, v + ncomp, values.begin() + static_cast<long long>(ncomp) * id);
        seen[id] = 1;
    };
    auto readCsv = [&](const std::string& csvPath) {
        std::ifstream csv(csvPath);
        if (!csv) throw std::runtime_error("cannot open restart field CSV: " + csvPath);
        std::string line;
        if (!std::getline(csv, line)) throw std::runtime_error("restart field CSV is empty: " + csvPath);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::ostringstream header;
        header << "global_id,x,y,z";
        for (int c = 0; c < ncomp; ++c) header << ",value" << c;
        if (line != header.str()) throw std::runtime_error("unexpected restart field CSV header in " + csvPath);
        std::size_t lineNo = 1;
        while (std::getline(csv, line)) {
            ++lineNo;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.empty()) continue;
            std::replace(line.begin(), line.end(), ',', ' ');
            std::istringstream row(line);
            long long id = -1;
            double x, y, z, v[9];
            bool ok = static_cast<bool>(row >> id >> x >> y >> z);
            for (int c = 0; ok && c < ncomp; ++c) ok = static_cast<bool>(row >> v[c]);
            if (!ok) throw std::runtime_error("invalid restart field row in " + csvPath + ":" + std::to_string(lineNo));
            std::string trailing;
            if (row >> trailing)
                throw std::runtime_error("unexpected data after restart field value in " + csvPath + ":" +
                                         std::to_string(lineNo));
            store(id, v, csvPath, lineNo);
        }
    };
    if (fs::is_directory(path)) {
        std::vector<std::string> files;
        for (const auto& e : fs::directory_iterator(path)) {
            if (!e.is_directory() || e.path().filename().string().rfind("rank-", 0) != 0) continue;
            const fs::path candidate = e.path() / (name + ".csv");
            if (fs::is_regular_file(candidate)) files.push_back(candidate.string());
        }
        std::sort(files.begin(), files.end());
        if (files.empty())
            throw std::runtime_error("no rank-sharded restart CSVs for field " + name + " under " + path);
        for (const auto& f : files) readCsv(f);
    } else if (fs::path(path).extension() == ".csv") {
        readCsv(path);
    } else {
        // "id value..." rows with # comments (field_reader.cpp:217-240).
        std::ifstream data(path);
        if (!data) throw std::runtime_error("cannot open nonuniform field data: " + path);
        std::string text;
        std::size_t lineNo = 0;
        while (std::getline(data, text)) {
            ++lineNo;
            const std::size_t comment = text.find('#');
            if (comment != std::string::npos) text.resize(comment);
            std::istringstream row(text);
            long long id = -1;
            if (!(row >> id)) continue;
            double v[9];
            bool ok = true;
            for (int c = 0; ok && c < ncomp; ++c) ok = static_cast<bool>(row >> v[c]);
            if (!ok) throw std::runtime_error("invalid field value in " + path + ":" + std::to_string(lineNo));
            std::string trailing;
            if (row >> trailing)
                throw std::runtime_error("unexpected data after field value in " + path + ":" + std::to_string(lineNo));
            store(id, v, path, lineNo);
        }
    }
    if (std::find(seen.begin(), seen.end(), 0) != seen.end())
        throw std::runtime_error("nonuniform field data is missing global cells: " + path);
    std::copy(values.begin(), values.end(), out);
}

}  // namespace bsim
